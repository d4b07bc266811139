// Device-scalar Lanczos building blocks for real (double) and complex
// (double2) Krylov vectors. Scalars (α_j, β_j, norms, overlap vectors) stay in
// device slots, so a Lanczos step runs without any host round trip; every
// pass streams whole columns of V (column-major, ld = n).
//
// Grid-wide sums are finished by the consuming kernel: the producer writes
// one partial per block and every block of the consumer adds them again in
// the same fixed order, so all blocks hold bit-identical totals and no
// in-kernel fence between blocks is needed (an agent-scope release writes back
// the whole L2 of an XCD, which costs more than the redundant adds from L2).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <type_traits>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"
#include "lanczos_kernels.hpp"

namespace qmbed {
namespace lk {

namespace {

constexpr int kT = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kRPT = 2;  // rows per thread in the overlap kernel
constexpr int kMaxCols = 512;

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

// block-wide sum; the result is valid in thread 0
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

// Σ partial[0..nb) in a fixed order; valid in thread 0
__device__ __forceinline__ double sum_partials(const double* partial, int nb, double* sh) {
  double t = 0;
  for (int b = threadIdx.x; b < nb; b += kT) t += partial[b];
  return block_sum(t, sh);
}

int nblocks(size_t n) { return int(std::min<size_t>(kMaxBlocks, std::max<size_t>(1, (n + kT * 4 - 1) / (kT * 4)))); }

// partial[block] = Re Σ conj(x) y over the block's rows
template <typename E>
__global__ __launch_bounds__(kT) void k_redot(const E* __restrict__ x, const E* __restrict__ y, size_t n,
                                              double* partial) {
  __shared__ double sh[kT / 64];
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) s += re_of(cmul(x[i], y[i]));
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[block] = Σ |y − λ x|² over the block's rows (λ from a device slot)
template <typename E>
__global__ __launch_bounds__(kT) void k_resid(const E* __restrict__ y, const E* __restrict__ x, const double* lam_slot,
                                              size_t n, double* partial) {
  __shared__ double sh[kT / 64];
  const double lam = *lam_slot;
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT)
    s += sq(sub(y[i], scale(x[i], lam)));
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// slot = Σ partial (its square root when root), one block
__global__ __launch_bounds__(kT) void k_sum(const double* partial, int nb, double* slot, int root) {
  __shared__ double sh[kT / 64];
  const double t = sum_partials(partial, nb, sh);
  if (threadIdx.x == 0) *slot = root ? sqrt(t) : t;
}

// partial[block] = Re Σ conj(v_j) w over the block's rows (α_j of a step)
template <typename E>
__device__ __forceinline__ void alpha_body(const E* __restrict__ vj, const E* __restrict__ w, size_t n,
                                           double* partial, const int* gate) {
  if (gate && *gate) return;
  __shared__ double sh[kT / 64];
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT)
    s += re_of(cmul(vj[i], w[i]));
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
template <typename E>
__global__ __launch_bounds__(kT) void k_alpha(const E* __restrict__ vj, const E* __restrict__ w, size_t n,
                                              double* partial, const int* gate) {
  alpha_body<E>(vj, w, n, partial, gate);
}

// α_j = Σ of the k_alpha partials (fixed order), once, into the step's α
// slot, where the overlap and subtraction blocks read it (one block; the
// blocks of those passes each summed the partials before)
__device__ __forceinline__ void alpha_fin_body(const double* partial, int nb, double* alpha, const int* gate) {
  if (gate && *gate) return;
  __shared__ double sh[kT / 64];
  const double a = sum_partials(partial, nb, sh);
  if (threadIdx.x == 0) *alpha = a;
}
__global__ __launch_bounds__(kT) void k_alpha_fin(const double* partial, int nb, double* alpha, const int* gate) {
  alpha_fin_body(partial, nb, alpha, gate);
}

// The three-term part of a first pass (QMBED src/solve/mod.rs:1040-1048):
// w₀ = w − α_j v_j − β_{j−1} v_{j−1}, α_j from its slot (k_alpha_fin).
// vj == nullptr: no three-term part.
template <typename E>
struct Three {
  const E* vj = nullptr;
  const E* vjm1 = nullptr;           // null at j = 0
  const double* beta_prev = nullptr;  // β_{j−1}
  const double* alpha = nullptr;      // α_j slot
};

// α_j and β_{j−1} into (alpha, beta) in every thread of the block
template <typename E>
__device__ __forceinline__ void three_coefs(const Three<E>& t, double& alpha, double& beta) {
  alpha = *t.alpha;
  beta = t.vjm1 ? *t.beta_prev : 0.0;
}

template <typename E>
__device__ __forceinline__ E three_apply(const Three<E>& t, E x, size_t i, double alpha, double beta) {
  x = sub(x, scale(t.vj[i], alpha));
  if (t.vjm1) x = sub(x, scale(t.vjm1[i], beta));
  return x;
}

// Overlaps of one classical Gram–Schmidt pass, per row block: for columns
// c < m, Σ_i conj(V[c][i]) w₀[i], and when mm == m + 1 also Σ_i |w₀[i]|² at
// c = m (‖w₀‖, the DGKS reference norm). w₀ is w after the three-term part
// (first pass) or w itself. Block (x, y) takes the row tiles x, x + gridDim.x,
// … of kT·RPT rows and the columns [y·CT, y·CT + CT); every load of a tile
// is unconditional (row and column indices are clamped, out-of-range rows
// carry w₀ = 0), so the CT·RPT loads of a thread are in flight together.
// Output: partial[x·ld + col_off + c].
// (bx, by, nbx: the block's row block, column group, and the number of row
// blocks, so that one launch can hold the blocks of several windows)
template <typename E, int CT, int RPT>
__device__ __forceinline__ void dots_body(const E* __restrict__ V, size_t n, int m, int mm, const E* __restrict__ w,
                                          E* __restrict__ partial, int ld, int col_off, const int* gate,
                                          const Three<E>& t, unsigned bx, unsigned by, unsigned nbx) {
  if (gate && *gate) return;
  __shared__ E sh[CT][kT / 64];
  double alpha = 0, beta = 0;
  if (t.vj) three_coefs(t, alpha, beta);
  const int c0 = by * CT;
  const int nc = min(CT, mm - c0);
  const size_t tiles = (n + size_t(kT * RPT) - 1) / size_t(kT * RPT);
  E acc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) acc[c] = zero<E>();
  // v_j and v_{j−1} are the columns m − 1 and m − 2 of V: when they lie in
  // this block's tile the three-term part takes the loaded values instead of
  // reading them again
  const int cj = m - 1 - c0, cjm1 = m - 2 - c0;
  const bool tile_j = t.vj && t.vj == V + size_t(m - 1) * n && cj >= 0 && cj < CT;
  const bool tile_jm1 = t.vjm1 && m >= 2 && t.vjm1 == V + size_t(m - 2) * n && cjm1 >= 0 && cjm1 < CT;
  for (size_t tt = bx; tt < tiles; tt += nbx) {
    const size_t r0 = tt * (kT * RPT) + threadIdx.x;
    size_t ic[RPT];
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      const size_t i = r0 + size_t(rr) * kT;
      ic[rr] = i < n ? i : n - 1;
    }
    E vv[CT][RPT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const size_t base = size_t(min(c0 + c, m - 1)) * n;
#pragma unroll
      for (int rr = 0; rr < RPT; ++rr) vv[c][rr] = V[base + ic[rr]];
    }
    E wv[RPT];
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
      E x = w[ic[rr]];
      if (t.vj) {
        E vj = zero<E>(), vm = zero<E>();
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          if (c == cj) vj = vv[c][rr];
          if (c == cjm1) vm = vv[c][rr];
        }
        if (!tile_j) vj = t.vj[ic[rr]];
        x = sub(x, scale(vj, alpha));
        if (t.vjm1) {
          if (!tile_jm1) vm = t.vjm1[ic[rr]];
          x = sub(x, scale(vm, beta));
        }
      }
      wv[rr] = r0 + size_t(rr) * kT < n ? x : zero<E>();
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const bool is_w = c0 + c >= m;
#pragma unroll
      for (int rr = 0; rr < RPT; ++rr) acc[c] = add(acc[c], cmul(is_w ? wv[rr] : vv[c][rr], wv[rr]));
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    E v = acc[c];
    for (int o = 32; o > 0; o >>= 1) v = add(v, shfl_down(v, o));
    if (lane == 0) sh[c][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    E s = zero<E>();
    for (int w2 = 0; w2 < kT / 64; ++w2) s = add(s, sh[threadIdx.x][w2]);
    partial[size_t(bx) * ld + col_off + c0 + threadIdx.x] = s;
  }
}
template <typename E, int CT, int RPT = kRPT>
__global__ __launch_bounds__(kT) void k_dots(const E* __restrict__ V, size_t n, int m, int mm, const E* __restrict__ w,
                                             E* __restrict__ partial, int ld, int col_off, const int* gate,
                                             Three<E> t) {
  dots_body<E, CT, RPT>(V, n, m, mm, w, partial, ld, col_off, gate, t, blockIdx.x, blockIdx.y, gridDim.x);
}

// h[c] = Σ_b partial[b·ld + c] over the rb row blocks of the overlap kernel,
// one block per column: thread-strided sums, then the waves' shuffle trees
// and their four totals in order (a fixed order).
template <typename E>
__device__ __forceinline__ void hred_body(const E* __restrict__ partial, int rb, int ld, int ncols, E* h,
                                          const int* gate) {
  if (gate && *gate) return;
  __shared__ E sh[kT / 64];
  const int c = blockIdx.x;
  if (c >= ncols) return;
  E s = zero<E>();
  for (int b = threadIdx.x; b < rb; b += kT) s = add(s, partial[size_t(b) * ld + c]);
  for (int o = 32; o > 0; o >>= 1) s = add(s, shfl_down(s, o));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    E t = zero<E>();
    for (int w = 0; w < kT / 64; ++w) t = add(t, sh[w]);
    h[c] = t;
  }
}
template <typename E>
__global__ __launch_bounds__(kT) void k_hred(const E* __restrict__ partial, int rb, int ld, int ncols, E* h,
                                             const int* gate) {
  hred_body<E>(partial, rb, ld, ncols, h, gate);
}

// One column chunk of the subtraction w ← w₀ − V h (hc: the chunk's
// overlaps). The three-term part is applied by the first chunk only, whose
// block 0 also stores α_j. With partial2 (last chunk) the per-block ‖w‖²
// partials of the result follow.
template <typename E>
__device__ __forceinline__ void sub_body(const E* __restrict__ V, size_t n, int mc, const E* __restrict__ hc,
                                         E* __restrict__ w, double* __restrict__ partial2, const int* gate,
                                         const Three<E>& t, unsigned bx, unsigned nbx) {
  if (gate && *gate) return;
  __shared__ E sh_h[kMaxCols];
  __shared__ double sh[kT / 64];
  double alpha = 0, beta = 0;
  if (t.vj) three_coefs(t, alpha, beta);
  for (int c = threadIdx.x; c < mc; c += kT) sh_h[c] = hc[c];
  __syncthreads();
  // v_j and v_{j−1} are normally the last two columns of this chunk: their
  // loads serve both the subtraction and the three-term part
  const bool last_j = t.vj && t.vj == V + size_t(mc - 1) * n;
  const bool last_jm1 = last_j && t.vjm1 && mc >= 2 && t.vjm1 == V + size_t(mc - 2) * n;
  const int cend = mc - (last_j ? (last_jm1 ? 2 : 1) : 0);
  double s = 0;
  bool pairs = false;
  if constexpr (std::is_same<E, double>::value)
    pairs = (n % 2 == 0) && ((reinterpret_cast<uintptr_t>(V) | reinterpret_cast<uintptr_t>(w)) % 16 == 0);
  if (pairs) {
    if constexpr (std::is_same<E, double>::value) {
      // two adjacent rows per thread, one 16-byte load per column: each row's
      // arithmetic is the single-row loop's below, term for term
      const size_t np = n / 2;
      const double2* __restrict__ V2 = reinterpret_cast<const double2*>(V);
      double2* __restrict__ w2 = reinterpret_cast<double2*>(w);
      for (size_t q = bx * size_t(kT) + threadIdx.x; q < np; q += size_t(nbx) * kT) {
        double a0 = 0.0, a1 = 0.0;
        int c = 0;
        for (; c + 4 <= cend; c += 4) {
          const double2 u0 = V2[size_t(c) * np + q], u1 = V2[size_t(c + 1) * np + q];
          const double2 u2 = V2[size_t(c + 2) * np + q], u3 = V2[size_t(c + 3) * np + q];
          a0 = add(a0, mul(u0.x, sh_h[c]));
          a1 = add(a1, mul(u0.y, sh_h[c]));
          a0 = add(a0, mul(u1.x, sh_h[c + 1]));
          a1 = add(a1, mul(u1.y, sh_h[c + 1]));
          a0 = add(a0, mul(u2.x, sh_h[c + 2]));
          a1 = add(a1, mul(u2.y, sh_h[c + 2]));
          a0 = add(a0, mul(u3.x, sh_h[c + 3]));
          a1 = add(a1, mul(u3.y, sh_h[c + 3]));
        }
        for (; c < cend; ++c) {
          const double2 u = V2[size_t(c) * np + q];
          a0 = add(a0, mul(u.x, sh_h[c]));
          a1 = add(a1, mul(u.y, sh_h[c]));
        }
        const double2 xw = w2[q];
        double x0 = xw.x, x1 = xw.y;
        const size_t i0 = 2 * q, i1 = 2 * q + 1;
        if (last_j) {
          const double2 vj = V2[size_t(mc - 1) * np + q];
          double2 vm = make_double2(0.0, 0.0);
          if (last_jm1) {
            vm = V2[size_t(mc - 2) * np + q];
            a0 = add(a0, mul(vm.x, sh_h[mc - 2]));
            a1 = add(a1, mul(vm.y, sh_h[mc - 2]));
          }
          a0 = add(a0, mul(vj.x, sh_h[mc - 1]));
          a1 = add(a1, mul(vj.y, sh_h[mc - 1]));
          x0 = sub(x0, scale(vj.x, alpha));
          x1 = sub(x1, scale(vj.y, alpha));
          if (t.vjm1) {
            x0 = sub(x0, scale(last_jm1 ? vm.x : t.vjm1[i0], beta));
            x1 = sub(x1, scale(last_jm1 ? vm.y : t.vjm1[i1], beta));
          }
        } else if (t.vj) {
          x0 = three_apply(t, x0, i0, alpha, beta);
          x1 = three_apply(t, x1, i1, alpha, beta);
        }
        const double v0 = sub(x0, a0), v1 = sub(x1, a1);
        w2[q] = make_double2(v0, v1);
        s += sq(v0);
        s += sq(v1);
      }
    }
  } else
  for (size_t i = bx * size_t(kT) + threadIdx.x; i < n; i += size_t(nbx) * kT) {
    E acc = zero<E>();
    int c = 0;
    for (; c + 4 <= cend; c += 4) {
      acc = add(acc, mul(V[size_t(c) * n + i], sh_h[c]));
      acc = add(acc, mul(V[size_t(c + 1) * n + i], sh_h[c + 1]));
      acc = add(acc, mul(V[size_t(c + 2) * n + i], sh_h[c + 2]));
      acc = add(acc, mul(V[size_t(c + 3) * n + i], sh_h[c + 3]));
    }
    for (; c < cend; ++c) acc = add(acc, mul(V[size_t(c) * n + i], sh_h[c]));
    E x = w[i];
    if (last_j) {
      const E vj = V[size_t(mc - 1) * n + i];
      E vm = zero<E>();
      if (last_jm1) {
        vm = V[size_t(mc - 2) * n + i];
        acc = add(acc, mul(vm, sh_h[mc - 2]));
      }
      acc = add(acc, mul(vj, sh_h[mc - 1]));
      x = sub(x, scale(vj, alpha));
      if (t.vjm1) x = sub(x, scale(last_jm1 ? vm : t.vjm1[i], beta));
    } else if (t.vj) {
      x = three_apply(t, x, i, alpha, beta);
    }
    const E v = sub(x, acc);
    w[i] = v;
    s += sq(v);
  }
  if (!partial2) return;
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial2[bx] = s;
}
template <typename E>
__global__ __launch_bounds__(kT) void k_sub(const E* __restrict__ V, size_t n, int mc, const E* __restrict__ hc,
                                            E* __restrict__ w, double* __restrict__ partial2, const int* gate,
                                            Three<E> t) {
  sub_body<E>(V, n, mc, hc, w, partial2, gate, t, blockIdx.x, gridDim.x);
}

// End of a Lanczos step. Every block forms ‖w‖² = Σ partial2 (fixed order).
// First pass (pass2 == 0): the DGKS decision — a second pass iff
// ‖w‖ <= η‖w₀‖ with ‖w₀‖ > eps (QMBED src/solve/mod.rs:1050-1073, η = 1/√2
// there), ‖w₀‖² = h[j+1] from the overlaps. A needed second pass sets
// *halt = j + 1 and leaves w as it is; every later gated launch is then a
// no-op until the host has run the second pass. Otherwise (and after a second
// pass, pass2 == 1, which clears *halt) β_j = ‖w‖ and out = w/β_j (+ FP32
// copy); out == nullptr writes β_j only. A first pass is skipped while *halt
// is set.
template <typename E, typename F>
__device__ __forceinline__ void finish_body(const E* __restrict__ w, size_t n, const double* __restrict__ partial2,
                                            int nb2, const E* __restrict__ h, int j, double eta, double* beta,
                                            int* halt, int* hist, E* out, F* out32, int pass2, const Interleave& il) {
  if (!pass2 && *halt) return;
  __shared__ double sh[kT / 64];
  __shared__ double s_after;
  const double t = sum_partials(partial2, nb2, sh);
  if (threadIdx.x == 0) s_after = t;
  __syncthreads();
  const double after2 = s_after;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!pass2) {
    const double nb = sqrt(re_of(h[j + 1])), na = sqrt(after2);
    const bool second = nb > 2.220446049250313e-16 && na <= eta * nb;
    if (lead) {
      *hist = second ? 1 : 0;
      if (second) *halt = j + 1;
    }
    if (second) return;
  } else if (lead) {
    *halt = 0;
  }
  const double b = sqrt(after2);
  if (lead) *beta = b;
  if (!out) return;
  const double inv = 1.0 / b;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const E v = scale(w[i], inv);
    out[i] = v;
    if (out32) {
      if constexpr (sizeof(E) == 8) out32[i] = float(re_of(v));
      else {
        out32[2 * i] = float(v.x);
        out32[2 * i + 1] = float(v.y);
      }
    }
    if constexpr (sizeof(E) == 8) {
      if (il.x64) il.x64[i * il.stride + il.lane] = re_of(v);
      if (il.x32) il.x32[i * il.stride + il.lane] = float(re_of(v));
    }
  }
}
template <typename E, typename F>
__global__ __launch_bounds__(kT) void k_finish(const E* __restrict__ w, size_t n, const double* __restrict__ partial2,
                                               int nb2, const E* __restrict__ h, int j, double eta, double* beta,
                                               int* halt, int* hist, E* out, F* out32, int pass2, Interleave il) {
  finish_body<E, F>(w, n, partial2, nb2, h, j, eta, beta, halt, hist, out, out32, pass2, il);
}

// Lockstep windows (real): one launch of a step phase for several windows,
// blockIdx.y = window, each block doing exactly what the single-window
// kernel's block does.
constexpr int kMaxJobs = 16;
struct AlphaJobs {
  const double* vj[kMaxJobs];
  const double* w[kMaxJobs];
  double* partial[kMaxJobs];
  const int* gate[kMaxJobs];
  double* alpha[kMaxJobs];
};
__global__ __launch_bounds__(kT) void k_alpha_multi(AlphaJobs J, size_t n) {
  const int q = blockIdx.y;
  alpha_body<double>(J.vj[q], J.w[q], n, J.partial[q], J.gate[q]);
}
__global__ __launch_bounds__(kT) void k_alpha_fin_multi(AlphaJobs J, int nb) {
  const int q = blockIdx.x;
  alpha_fin_body(J.partial[q], nb, J.alpha[q], J.gate[q]);
}
struct HredJobs {
  const double* partial[kMaxJobs];
  int rb[kMaxJobs], ld[kMaxJobs], ncols[kMaxJobs];
  double* h[kMaxJobs];
  const int* gate[kMaxJobs];
};
__global__ __launch_bounds__(kT) void k_hred_multi(HredJobs J) {
  const int q = blockIdx.y;
  hred_body<double>(J.partial[q], J.rb[q], J.ld[q], J.ncols[q], J.h[q], J.gate[q]);
}
struct FinishJobs {
  const double* w[kMaxJobs];
  const double* partial2[kMaxJobs];
  const double* h[kMaxJobs];
  int j[kMaxJobs];
  double* beta[kMaxJobs];
  int* halt[kMaxJobs];
  int* hist[kMaxJobs];
  double* out[kMaxJobs];
};
__global__ __launch_bounds__(kT) void k_finish_multi(FinishJobs J, size_t n, int nb2, double eta) {
  const int q = blockIdx.y;
  const Interleave none;
  finish_body<double, float>(J.w[q], n, J.partial2[q], nb2, J.h[q], J.j[q], eta, J.beta[q], J.halt[q], J.hist[q],
                             J.out[q], static_cast<float*>(nullptr), 0, none);
}

// out = w / sqrt(*nrm2); beta_slot (optional) = sqrt(*nrm2); optional FP32
// copy; out == nullptr writes beta_slot only
template <typename E, typename F>
__global__ __launch_bounds__(kT) void k_normalize(const E* w, const double* nrm2, size_t n, E* out, F* out32,
                                                  double* beta_slot) {
  const double b = sqrt(*nrm2);
  const double inv = 1.0 / b;
  if (beta_slot && blockIdx.x == 0 && threadIdx.x == 0) *beta_slot = b;
  if (!out) return;
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
__global__ __launch_bounds__(kT) void k_gemm_vc(const E* __restrict__ V, size_t n, int m, const E* __restrict__ C,
                                                E* Y) {
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

// Tall-skinny Aᵀ B, one 32 × 32 output tile per blockIdx.y: the block's
// chunks of 64 rows of both operands are staged in LDS (coalesced column
// loads; absent rows and columns read as zero); four groups of 64 threads
// take every fourth row of a chunk, each thread a 4 × 4 register tile of the
// output; the groups' sums meet in LDS and the block writes its partial
// tile to partial[(blockIdx.y · gridDim.x + blockIdx.x) · 1024 + i·32 + j].
constexpr int kAtbTile = 32, kAtbRows = 64;
__global__ __launch_bounds__(kT) void k_atb(const double* __restrict__ A, size_t n, int p, const double* __restrict__ B,
                                            int q, int tiles_j, double* __restrict__ partial) {
  // the operand tiles, then (after the last chunk) the groups' sums
  constexpr int kLd = kAtbTile + 1;
  __shared__ double sm[2 * kAtbRows * kLd];
  static_assert(2 * kAtbRows * kLd >= 4 * 16 * 64, "LDS of the group sums");
  double* sa = sm;
  double* sb = sm + kAtbRows * kLd;
  const int ti = blockIdx.y / tiles_j, tj = blockIdx.y % tiles_j;
  const int t = threadIdx.x, grp = t >> 6, lane = t & 63;
  const int i0 = (lane >> 3) * 4, j0 = (lane & 7) * 4;
  double acc[4][4] = {};
  const size_t chunks = (n + kAtbRows - 1) / kAtbRows;
  for (size_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const size_t r0 = ch * kAtbRows;
    for (int e = t; e < kAtbRows * kAtbTile; e += kT) {
      const int rr = e % kAtbRows, cc = e / kAtbRows;
      const size_t r = r0 + size_t(rr);
      const int ca = ti * kAtbTile + cc, cb = tj * kAtbTile + cc;
      sa[rr * kLd + cc] = (r < n && ca < p) ? A[size_t(ca) * n + r] : 0.0;
      sb[rr * kLd + cc] = (r < n && cb < q) ? B[size_t(cb) * n + r] : 0.0;
    }
    __syncthreads();
    for (int rr = grp; rr < kAtbRows; rr += 4) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = sa[rr * kLd + i0 + u];
        b[u] = sb[rr * kLd + j0 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
    __syncthreads();
  }
  double* red = sm;  // [grp][16][64]; the loop ended with a barrier
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) red[(grp * 16 + u * 4 + v) * 64 + lane] = acc[u][v];
  __syncthreads();
  if (grp == 0) {
    double* out = partial + (size_t(blockIdx.y) * gridDim.x + blockIdx.x) * (kAtbTile * kAtbTile);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int k = u * 4 + v;
        out[(i0 + u) * kAtbTile + j0 + v] =
            ((red[k * 64 + lane] + red[(16 + k) * 64 + lane]) + red[(32 + k) * 64 + lane]) + red[(48 + k) * 64 + lane];
      }
  }
}

// C[i + p·j] = Σ_b partial of tile (i/32, j/32), block b (in block order)
__global__ __launch_bounds__(kT) void k_atb_reduce(const double* __restrict__ partial, int blocks, int p, int q,
                                                   int tiles_j, double* __restrict__ C) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= p * q) return;
  const int i = e % p, j = e / p;
  const int tile = (i / kAtbTile) * tiles_j + j / kAtbTile;
  const int k = (i % kAtbTile) * kAtbTile + j % kAtbTile;
  double s = 0;
  for (int b = 0; b < blocks; ++b) s += partial[(size_t(tile) * blocks + b) * (kAtbTile * kAtbTile) + k];
  C[e] = s;
}

template <typename E>
__global__ __launch_bounds__(kT) void k_to32(const E* x, size_t n, float* y) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    if constexpr (sizeof(E) == 8) y[i] = float(re_of(x[i]));
    else {
      y[2 * i] = float(x[i].x);
      y[2 * i + 1] = float(x[i].y);
    }
  }
}
template <typename E>
__global__ __launch_bounds__(kT) void k_from32(const float* y, size_t n, E* x, const int* gate) {
  if (gate && *gate) return;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    if constexpr (sizeof(E) == 8) x[i] = double(y[i]);
    else x[i] = make_double2(double(y[2 * i]), double(y[2 * i + 1]));
  }
}

// deterministic_start (QMBED src/solve/mod.rs:729-742)
__global__ void k_det_start(double* out, size_t n, uint64_t seed, int complex_pairs) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double v = deterministic_entry(uint64_t(i), seed);
    if (complex_pairs) {
      out[2 * i] = v;
      out[2 * i + 1] = 0.0;
    } else {
      out[i] = v;
    }
  }
}

// column tile of the overlap kernel: 16 real or 8 complex columns
template <typename E>
constexpr int dots_tile() {
  return sizeof(E) == 8 ? 16 : 8;
}

}  // namespace

Workspace::Workspace() : partial(64 * 1024), partial2(kMaxBlocks), partial3(kMaxBlocks), slots(4096) {}

template <typename E>
static const E* cE(const void* p) {
  return static_cast<const E*>(p);
}
template <typename E>
static E* mE(void* p) {
  return static_cast<E*>(p);
}

#define QMBED_DISPATCH(cplx_flag, ...) \
  do {                                 \
    if (cplx_flag) {                   \
      using E = double2;               \
      __VA_ARGS__;                     \
    } else {                           \
      using E = double;                \
      __VA_ARGS__;                     \
    }                                  \
  } while (0)

void redot_to_slot(bool c, const void* x, const void* y, size_t n, double* slot, Workspace& ws, hipStream_t s,
                   bool root) {
  const int nb = nblocks(n);
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_redot<E>, dim3(nb), dim3(kT), 0, s, cE<E>(x), cE<E>(y), n,
                                       ws.partial2.data()));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(kT), 0, s, ws.partial2.data(), nb, slot, root ? 1 : 0);
}

void nrm2sq_to_slot(bool c, const void* x, size_t n, double* slot, Workspace& ws, hipStream_t s) {
  redot_to_slot(c, x, x, n, slot, ws, s, false);
}

void normalize(bool c, const void* w, const double* nrm2_slot, size_t n, void* out, float* out32, double* beta_slot,
               hipStream_t s) {
  QMBED_DISPATCH(c, hipLaunchKernelGGL((k_normalize<E, float>), dim3(nblocks(n)), dim3(kT), 0, s, cE<E>(w), nrm2_slot,
                                       n, mE<E>(out), out32, beta_slot));
}

void gemm_vc(bool c, const void* V, size_t n, int m, const void* d_C, int k, void* Y, hipStream_t s) {
  const size_t esz = c ? 16 : 8;
  // chunks of up to 16 columns: V is read once per chunk, and every column's
  // sum runs in the same order whatever the chunk
  for (int q0 = 0; q0 < k;) {
    const int left = k - q0;
    int kk = left >= 16 ? 16 : left >= 8 ? 8 : std::min(4, left);
    if (size_t(m) * kk * esz > 64 * 1024) kk = std::min(4, left);
    size_t sm = size_t(m) * kk * esz;
    if (sm > 64 * 1024) throw unsupported_backend("Ritz assembly window exceeds the LDS coefficient budget");
    QMBED_DISPATCH(c, {
      const E* C = cE<E>(d_C) + size_t(q0) * m;
      E* Yq = mE<E>(Y) + size_t(q0) * n;
      switch (kk) {
        case 1: hipLaunchKernelGGL((k_gemm_vc<E, 1>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
        case 2: hipLaunchKernelGGL((k_gemm_vc<E, 2>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
        case 3: hipLaunchKernelGGL((k_gemm_vc<E, 3>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
        case 4: hipLaunchKernelGGL((k_gemm_vc<E, 4>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
        case 8: hipLaunchKernelGGL((k_gemm_vc<E, 8>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
        default: hipLaunchKernelGGL((k_gemm_vc<E, 16>), dim3(nblocks(n)), dim3(kT), sm, s, cE<E>(V), n, m, C, Yq); break;
      }
    });
    q0 += kk;
  }
}

// Starts of lockstep windows: the squared norms of nv vectors (the per-block
// partials and their sums of nrm2sq_to_slot, for every vector at once), then
// out_q = x_q / ‖x_q‖ (as normalize) with the window's flag area cleared.
struct NormJobs {
  const double* x[kMaxJobs];
};
__global__ __launch_bounds__(kT) void k_nrm2_multi(NormJobs J, size_t n, double* partial) {
  __shared__ double sh[kT / 64];
  const double* x = J.x[blockIdx.y];
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) s += x[i] * x[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}
__global__ __launch_bounds__(kT) void k_sum_multi(const double* partial, int nb, double* slots) {
  __shared__ double sh[kT / 64];
  const double t = sum_partials(partial + size_t(blockIdx.x) * nb, nb, sh);
  if (threadIdx.x == 0) slots[blockIdx.x] = t;
}
struct StartJobs {
  const double* x[kMaxJobs];
  double* out[kMaxJobs];
  double* zero[kMaxJobs];
  int nzero[kMaxJobs];
};
__global__ __launch_bounds__(kT) void k_start_multi(StartJobs J, const double* nrm2, size_t n) {
  const int q = blockIdx.y;
  const double inv = 1.0 / sqrt(nrm2[q]);
  const double* x = J.x[q];
  double* out = J.out[q];
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) out[i] = x[i] * inv;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < J.nzero[q]; i += kT) J.zero[q][i] = 0.0;
}

void start_multi(const StartJob* jobs, int nj, size_t n, Workspace& ws, hipStream_t s) {
  if (nj < 1 || nj > kMaxJobs) throw internal_state("lockstep starts of 1 to 16 windows");
  const int nb = nblocks(n);
  if (ws.partial.size() < size_t(nj) * nb || ws.slots.size() < size_t(nj)) throw internal_state("start workspace too small");
  NormJobs N{};
  StartJobs S{};
  for (int q = 0; q < nj; ++q) {
    N.x[q] = jobs[q].x;
    S.x[q] = jobs[q].x;
    S.out[q] = jobs[q].out;
    S.zero[q] = jobs[q].zero;
    S.nzero[q] = jobs[q].nzero;
  }
  hipLaunchKernelGGL(k_nrm2_multi, dim3(nb, nj), dim3(kT), 0, s, N, n, ws.partial.data());
  hipLaunchKernelGGL(k_sum_multi, dim3(nj), dim3(kT), 0, s, ws.partial.data(), nb, ws.slots.data());
  hipLaunchKernelGGL(k_start_multi, dim3(nblocks(n), nj), dim3(kT), 0, s, S, ws.slots.data(), n);
  QMBED_HIP(hipGetLastError());
}

// the launch shape of gemm_atb: tiles of the output and row blocks, about
// 512 blocks in all, at least one row chunk each; partial doubles needed
static void atb_shape(size_t n, int p, int q, int& tiles_j, int& tiles, int& bx, size_t& need) {
  const int tiles_i = (p + kAtbTile - 1) / kAtbTile;
  tiles_j = (q + kAtbTile - 1) / kAtbTile;
  tiles = tiles_i * tiles_j;
  const size_t chunks = std::max<size_t>(1, (n + kAtbRows - 1) / kAtbRows);
  bx = int(std::min<size_t>(chunks, size_t(std::max(1, 512 / tiles))));
  need = size_t(tiles) * size_t(bx) * kAtbTile * kAtbTile;
}

// every block writes its whole partial tile, so `partial` needs no clearing
static void atb_launch(const double* A, size_t n, int p, const double* B, int q, double* C, double* partial,
                       int tiles_j, int tiles, int bx, hipStream_t s) {
  hipLaunchKernelGGL(k_atb, dim3(bx, tiles), dim3(kT), 0, s, A, n, p, B, q, tiles_j, partial);
  hipLaunchKernelGGL(k_atb_reduce, dim3((p * q + kT - 1) / kT), dim3(kT), 0, s, partial, bx, p, q, tiles_j, C);
  QMBED_HIP(hipGetLastError());
}

void gemm_atb(const double* A, size_t n, int p, const double* B, int q, double* C, DeviceArray<double>& partial,
              hipStream_t s) {
  if (p <= 0 || q <= 0) return;
  int tiles_j, tiles, bx;
  size_t need;
  atb_shape(n, p, q, tiles_j, tiles, bx, need);
  if (partial.size() < need) partial.resize(need);
  atb_launch(A, n, p, B, q, C, partial.data(), tiles_j, tiles, bx, s);
}

// The allocation-free variant's scratch: the partial products of
// gemm_atb_scratch, one copy per device (held until the process ends). Host
// code uses the device address hipGetSymbolAddress gives for it on each
// device, never the array's name.
constexpr size_t kAtbScratchDoubles = size_t(1) << 19;  // 4 MiB
static __device__ __align__(256) double atb_scratch[kAtbScratchDoubles];

size_t atb_scratch_capacity() { return kAtbScratchDoubles; }

// the current device's address of atb_scratch (looked up once per device)
static double* atb_scratch_ptr() {
  static std::mutex mu;
  static std::vector<double*> ptrs;
  int dev = 0;
  QMBED_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  if (size_t(dev) >= ptrs.size()) ptrs.resize(size_t(dev) + 1, nullptr);
  if (!ptrs[size_t(dev)]) {
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(atb_scratch));
    if (e != hipSuccess || !p) hip_check(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(atb_scratch)");
    ptrs[size_t(dev)] = p;
  }
  return ptrs[size_t(dev)];
}

bool gemm_atb_scratch(const double* A, size_t n, int p, const double* B, int q, double* C, hipStream_t s) {
  if (p <= 0 || q <= 0) return true;
  int tiles_j, tiles, bx;
  size_t need;
  atb_shape(n, p, q, tiles_j, tiles, bx, need);
  if (need > kAtbScratchDoubles) return false;
  atb_launch(A, n, p, B, q, C, atb_scratch_ptr(), tiles_j, tiles, bx, s);
  return true;
}

void to_f32(bool c, const void* x, size_t n, float* y, hipStream_t s) {
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_to32<E>, dim3(nblocks(n)), dim3(kT), 0, s, cE<E>(x), n, y));
}
void from_f32(bool c, const float* y, size_t n, void* x, hipStream_t s, const int* gate) {
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_from32<E>, dim3(nblocks(n)), dim3(kT), 0, s, y, n, mE<E>(x), gate));
}

void residual_to_slot(bool c, const void* y, const void* x, const double* lam_slot, size_t n, double* slot,
                      Workspace& ws, hipStream_t s) {
  const int nb = nblocks(n);
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_resid<E>, dim3(nb), dim3(kT), 0, s, cE<E>(y), cE<E>(x), lam_slot, n,
                                       ws.partial2.data()));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(kT), 0, s, ws.partial2.data(), nb, slot, 1);
}

// row blocks of the overlap kernel: about 1024 blocks in all, at most 256,
// never more than there are row tiles
// Row blocks of the overlap kernel: about 1024 blocks in all (four per CU,
// so the column loads of sixteen waves per CU are in flight), at least 16
// per column group, at most one per row tile.
static int dots_row_blocks(size_t n, int groups, int rpt) {
  const size_t tiles = (n + size_t(kT * rpt) - 1) / size_t(kT * rpt);
  const int want = std::max(16, 1024 / std::max(groups, 1));
  return int(std::min<size_t>(tiles, size_t(want)));
}

// Columns per block of the overlap kernel for a chunk of mm columns: the
// smallest tile that holds them (4, 8, then dots_tile), so short windows
// issue no loads for absent columns.
template <typename E>
static int dots_ct(int mm) {
  if (dots_tile<E>() == 8) return mm <= 4 ? 4 : 8;
  return mm <= 4 ? 4 : mm <= 8 ? 8 : mm <= 12 ? 12 : 16;
}

// rows per thread of the overlap kernel (4 for the narrow tiles measured
// slower: fewer blocks in flight)
static int dots_rpt(int) { return kRPT; }

template <typename E>
static int chunk_groups(int mm) {
  const int ct = dots_ct<E>(mm);
  return (mm + ct - 1) / ct;
}

// the layout shared by overlaps and subtract: rb from the widest chunk
template <typename E>
static void pass_layout(size_t n, int m, bool with_ww, int& rb, int& ld) {
  const int widest = std::min(m, kMaxCols) + (with_ww ? 1 : 0);
  rb = dots_row_blocks(n, chunk_groups<E>(widest), dots_rpt(dots_ct<E>(widest)));
  ld = m + 1;
}

template <typename E>
static Three<E> three_of(const ThreeTerm* tt, Workspace& ws, size_t n) {
  Three<E> t;
  if (!tt) return t;
  t.vj = cE<E>(tt->vj);
  t.vjm1 = cE<E>(tt->vjm1);
  t.beta_prev = tt->beta_prev;
  t.alpha = tt->alpha;
  (void)ws;
  (void)n;
  return t;
}

void lanczos_alpha(bool c, const void* vj, const void* w, size_t n, Workspace& ws, hipStream_t s, const int* gate,
                   double* alpha) {
  QMBED_DISPATCH(c, hipLaunchKernelGGL(k_alpha<E>, dim3(nblocks(n)), dim3(kT), 0, s, cE<E>(vj), cE<E>(w), n,
                                       ws.partial3.data(), gate));
  hipLaunchKernelGGL(k_alpha_fin, dim3(1), dim3(kT), 0, s, ws.partial3.data(), nblocks(n), alpha, gate);
}

void lanczos_overlaps(bool c, const void* V, size_t n, int m, const void* w, Workspace& ws, hipStream_t s,
                      bool with_ww, const int* gate, const ThreeTerm* three) {
  QMBED_DISPATCH(c, {
    int rb, ld;
    pass_layout<E>(n, m, with_ww, rb, ld);
    const size_t need = size_t(rb) * size_t(ld) * (sizeof(E) / 8);
    if (ws.partial.size() < need) ws.partial.resize(need * 2);
    const Three<E> t = three_of<E>(three, ws, n);
    for (int c0 = 0; c0 < m; c0 += kMaxCols) {
      const int mc = std::min(kMaxCols, m - c0);
      const bool last = c0 + mc >= m;
      const int mm = mc + ((last && with_ww) ? 1 : 0);
      const dim3 grid(rb, chunk_groups<E>(mm));
      const E* Vc = cE<E>(V) + size_t(c0) * n;
      switch (dots_ct<E>(mm)) {
        case 4:
          hipLaunchKernelGGL((k_dots<E, 4>), grid, dim3(kT), 0, s, Vc, n, mc, mm, cE<E>(w),
                             mE<E>(ws.partial.data()), ld, c0, gate, t);
          break;
        case 8:
          hipLaunchKernelGGL((k_dots<E, 8>), grid, dim3(kT), 0, s, Vc, n, mc, mm, cE<E>(w),
                             mE<E>(ws.partial.data()), ld, c0, gate, t);
          break;
        case 12:
          hipLaunchKernelGGL((k_dots<E, 12>), grid, dim3(kT), 0, s, Vc, n, mc, mm, cE<E>(w), mE<E>(ws.partial.data()),
                             ld, c0, gate, t);
          break;
        default:
          hipLaunchKernelGGL((k_dots<E, dots_tile<E>()>), grid, dim3(kT), 0, s, Vc, n, mc, mm, cE<E>(w),
                             mE<E>(ws.partial.data()), ld, c0, gate, t);
          break;
      }
    }
  });
}

// blocks of the subtract kernel: one row per thread up to 1024 blocks (its
// per-block norm partials are re-added by every block of the finishing
// kernel)
static int sub_blocks(size_t n) {
  return int(std::min<size_t>(kMaxBlocks, std::max<size_t>(1, (n + kT - 1) / kT)));
}

void lanczos_subtract(bool c, const void* V, size_t n, int m, void* w, void* d_h, Workspace& ws, hipStream_t s,
                      bool with_ww, const int* gate, const ThreeTerm* three, bool reduce) {
  const int nb = sub_blocks(n);
  QMBED_DISPATCH(c, {
    int rb, ld;
    pass_layout<E>(n, m, with_ww, rb, ld);
    const int ncols = m + (with_ww ? 1 : 0);
    if (reduce && ncols > 0)
      hipLaunchKernelGGL(k_hred<E>, dim3(ncols), dim3(kT), 0, s, cE<E>(ws.partial.data()),
                       rb, ld, ncols, mE<E>(d_h), gate);
    const Three<E> t = three_of<E>(three, ws, n);
    for (int c0 = 0; c0 < m; c0 += kMaxCols) {
      const int mc = std::min(kMaxCols, m - c0);
      const bool last = c0 + mc >= m;
      hipLaunchKernelGGL(k_sub<E>, dim3(nb), dim3(kT), 0, s, cE<E>(V) + size_t(c0) * n, n, mc, cE<E>(d_h) + c0,
                         mE<E>(w), last ? ws.partial2.data() : nullptr, gate, c0 == 0 ? t : Three<E>());
    }
  });
}

// The overlap and subtraction passes of several lockstep windows (real, one
// column chunk each, first passes) in one launch each: blockIdx.z = window,
// and each window's blocks compute exactly what its own launch would.
struct DotsJobs {
  const double* V[kMaxJobs];
  const double* w[kMaxJobs];
  double* partial[kMaxJobs];
  const int* gate[kMaxJobs];
  Three<double> t[kMaxJobs];
  int m[kMaxJobs], rb[kMaxJobs], groups[kMaxJobs], ld[kMaxJobs];
};
template <int CT>
__global__ __launch_bounds__(kT) void k_dots_multi(DotsJobs J, size_t n) {
  const int q = blockIdx.z;
  if (int(blockIdx.x) >= J.rb[q] || int(blockIdx.y) >= J.groups[q]) return;
  dots_body<double, CT, kRPT>(J.V[q], n, J.m[q], J.m[q] + 1, J.w[q], J.partial[q], J.ld[q], 0, J.gate[q], J.t[q],
                              blockIdx.x, blockIdx.y, unsigned(J.rb[q]));
}
struct SubJobs {
  const double* V[kMaxJobs];
  const double* h[kMaxJobs];
  double* w[kMaxJobs];
  double* partial2[kMaxJobs];
  const int* gate[kMaxJobs];
  Three<double> t[kMaxJobs];
  int m[kMaxJobs];
};
__global__ __launch_bounds__(kT) void k_sub_multi(SubJobs J, size_t n) {
  const int q = blockIdx.y;
  sub_body<double>(J.V[q], n, J.m[q], J.h[q], J.w[q], J.partial2[q], J.gate[q], J.t[q], blockIdx.x, gridDim.x);
}

void lanczos_overlaps_multi(const PassJob* jobs, int nj, size_t n, hipStream_t s) {
  if (nj < 1 || nj > kMaxJobs) throw internal_state("lockstep phase launch of 1 to 16 windows");
  // one launch per column tile in use (the windows of a batch mostly share one)
  for (int ct : {4, 8, 12, 16}) {
    DotsJobs J{};
    int k = 0, rbmax = 0, gmax = 0;
    for (int q = 0; q < nj; ++q) {
      const PassJob& jb = jobs[q];
      if (jb.m < 1 || jb.m > kMaxCols) throw internal_state("lockstep overlaps of a window beyond one column chunk");
      if (dots_ct<double>(jb.m + 1) != ct) continue;
      int rb, ld;
      pass_layout<double>(n, jb.m, true, rb, ld);
      const size_t need = size_t(rb) * size_t(ld);
      if (jb.ws->partial.size() < need) jb.ws->partial.resize(need * 2);
      J.V[k] = jb.V;
      J.w[k] = jb.w;
      J.partial[k] = jb.ws->partial.data();
      J.gate[k] = jb.gate;
      J.t[k] = three_of<double>(&jb.three, *jb.ws, n);
      J.m[k] = jb.m;
      J.rb[k] = rb;
      J.groups[k] = chunk_groups<double>(jb.m + 1);
      J.ld[k] = ld;
      rbmax = std::max(rbmax, rb);
      gmax = std::max(gmax, J.groups[k]);
      ++k;
    }
    if (k == 0) continue;
    const dim3 grid{unsigned(rbmax), unsigned(gmax), unsigned(k)};
    // algorithmic bytes: every window's columns and w once
    double bytes = 0;
    for (int q = 0; q < k; ++q) bytes += double(J.m[q] + 1) * 8.0 * double(n);
    char kname[32];
    std::snprintf(kname, sizeof kname, "lk::k_dots_multi<%d>", ct);
    const int timed = launch_timer_begin(kname, s, bytes);
    switch (ct) {
      case 4: hipLaunchKernelGGL(k_dots_multi<4>, grid, dim3(kT), 0, s, J, n); break;
      case 8: hipLaunchKernelGGL(k_dots_multi<8>, grid, dim3(kT), 0, s, J, n); break;
      case 12: hipLaunchKernelGGL(k_dots_multi<12>, grid, dim3(kT), 0, s, J, n); break;
      default: hipLaunchKernelGGL(k_dots_multi<16>, grid, dim3(kT), 0, s, J, n); break;
    }
    QMBED_HIP(hipGetLastError());
    launch_timer_end(timed, s);
  }
}

void lanczos_subtract_multi(const PassJob* jobs, int nj, size_t n, hipStream_t s) {
  if (nj < 1 || nj > kMaxJobs) throw internal_state("lockstep phase launch of 1 to 16 windows");
  SubJobs J{};
  for (int q = 0; q < nj; ++q) {
    const PassJob& jb = jobs[q];
    if (jb.m < 1 || jb.m > kMaxCols) throw internal_state("lockstep subtraction of a window beyond one column chunk");
    J.V[q] = jb.V;
    J.h[q] = jb.h;
    J.w[q] = jb.w;
    J.partial2[q] = jb.ws->partial2.data();
    J.gate[q] = jb.gate;
    J.t[q] = three_of<double>(&jb.three, *jb.ws, n);
    J.m[q] = jb.m;
  }
  // algorithmic bytes: every window's columns read once, w read and written
  double bytes = 0;
  for (int q = 0; q < nj; ++q) bytes += double(J.m[q] + 2) * 8.0 * double(n);
  const int timed = launch_timer_begin("lk::k_sub_multi", s, bytes);
  hipLaunchKernelGGL(k_sub_multi, dim3(unsigned(sub_blocks(n)), unsigned(nj)), dim3(kT), 0, s, J, n);
  QMBED_HIP(hipGetLastError());
  launch_timer_end(timed, s);
}

void lanczos_alpha_multi(const AlphaJob* jobs, int nj, size_t n, hipStream_t s) {
  if (nj < 1 || nj > kMaxJobs) throw internal_state("lockstep phase launch of 1 to 16 windows");
  AlphaJobs J{};
  for (int q = 0; q < nj; ++q) {
    J.vj[q] = jobs[q].vj;
    J.w[q] = jobs[q].w;
    J.partial[q] = jobs[q].ws->partial3.data();
    J.gate[q] = jobs[q].gate;
    J.alpha[q] = jobs[q].alpha;
  }
  hipLaunchKernelGGL(k_alpha_multi, dim3(nblocks(n), nj), dim3(kT), 0, s, J, n);
  QMBED_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_alpha_fin_multi, dim3(nj), dim3(kT), 0, s, J, nblocks(n));
  QMBED_HIP(hipGetLastError());
}

void lanczos_hred_multi(const HredJob* jobs, int nj, size_t n, hipStream_t s) {
  if (nj < 1 || nj > kMaxJobs) throw internal_state("lockstep phase launch of 1 to 16 windows");
  HredJobs J{};
  int widest = 1;
  for (int q = 0; q < nj; ++q) {
    int rb, ld;
    pass_layout<double>(n, jobs[q].m, true, rb, ld);
    J.partial[q] = jobs[q].ws->partial.data();
    J.rb[q] = rb;
    J.ld[q] = ld;
    J.ncols[q] = jobs[q].m + 1;
    J.h[q] = jobs[q].h;
    J.gate[q] = jobs[q].gate;
    widest = std::max(widest, jobs[q].m + 1);
  }
  hipLaunchKernelGGL(k_hred_multi, dim3(widest, nj), dim3(kT), 0, s, J);
  QMBED_HIP(hipGetLastError());
}

void lanczos_finish_multi(const FinishJob* jobs, int nj, size_t n, double eta, hipStream_t s) {
  if (nj < 1 || nj > kMaxJobs) throw internal_state("lockstep phase launch of 1 to 16 windows");
  FinishJobs J{};
  for (int q = 0; q < nj; ++q) {
    J.w[q] = jobs[q].w;
    J.partial2[q] = jobs[q].ws->partial2.data();
    J.h[q] = jobs[q].h;
    J.j[q] = jobs[q].j;
    J.beta[q] = jobs[q].beta;
    J.halt[q] = jobs[q].halt;
    J.hist[q] = jobs[q].hist;
    J.out[q] = jobs[q].out;
  }
  hipLaunchKernelGGL(k_finish_multi, dim3(nblocks(n), nj), dim3(kT), 0, s, J, n, sub_blocks(n), eta);
  QMBED_HIP(hipGetLastError());
}

struct CopyJobs {
  CopyJob j[kMaxJobs];
};
// block q copies job q
__global__ __launch_bounds__(kT) void k_copy_multi(CopyJobs J) {
  const CopyJob jb = J.j[blockIdx.x];
  for (int i = threadIdx.x; i < jb.count; i += kT) jb.dst[i] = jb.src[i];
}

void copy_multi(const CopyJob* jobs, int nj, hipStream_t s) {
  if (nj < 1 || nj > kMaxJobs) throw internal_state("batched copy of 1 to 16 arrays");
  CopyJobs J{};
  for (int q = 0; q < nj; ++q) {
    if (jobs[q].count < 0 || !jobs[q].src || !jobs[q].dst) throw internal_state("batched copy job without its arrays");
    J.j[q] = jobs[q];
  }
  hipLaunchKernelGGL(k_copy_multi, dim3(nj), dim3(kT), 0, s, J);
  QMBED_HIP(hipGetLastError());
}

void lanczos_finish(bool c, const void* w, size_t n, const void* d_h, int j, double eta, double* beta, int* halt,
                    int* hist, void* out, float* out32, Workspace& ws, hipStream_t s, bool second_pass,
                    const Interleave* il) {
  const Interleave none;
  if (il && c) throw internal_state("interleaved copies are for real vectors");
  QMBED_DISPATCH(c, hipLaunchKernelGGL((k_finish<E, float>), dim3(nblocks(n)), dim3(kT), 0, s, cE<E>(w), n,
                                       ws.partial2.data(), sub_blocks(n), cE<E>(d_h), j, eta, beta, halt, hist,
                                       mE<E>(out), out32, second_pass ? 1 : 0, il ? *il : none));
}

__global__ __launch_bounds__(kT) void k_scatter_lane(const double* __restrict__ x, size_t n, Interleave il) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    if (il.x64) il.x64[i * il.stride + il.lane] = x[i];
    if (il.x32) il.x32[i * il.stride + il.lane] = float(x[i]);
  }
}

struct Lanes16 {
  const double* src[16];
};
// one row per thread: the W values of a row are gathered from the W
// columns (coalesced reads) and stored side by side as whole vectors
template <typename T, int W>
__global__ __launch_bounds__(kT) void k_interleave(Lanes16 l, size_t n, T* __restrict__ x) {
  struct alignas(sizeof(T) * W) Row {
    T v[W];
  };
  Row* out = reinterpret_cast<Row*>(x);
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    Row r;
#pragma unroll
    for (int b = 0; b < W; ++b) r.v[b] = l.src[b] ? T(l.src[b][i]) : T(0);
    out[i] = r;
  }
}

void interleave(const double* const* src, size_t n, int width, double* x64, float* x32, hipStream_t s) {
  if (!(width == 1 || width == 2 || width == 4 || width == 8 || width == 16))
    throw internal_state("interleave width must be 1, 2, 4, 8 or 16");
  if (!x64 == !x32) throw internal_state("interleave writes exactly one of the FP64 and FP32 lanes");
  Lanes16 l{};
  for (int b = 0; b < width; ++b) l.src[b] = src[b];
  // one row per thread (the column loads of more waves in flight)
  const dim3 g(unsigned(std::max<size_t>(1, (n + kT - 1) / kT))), t(kT);
#define QMBED_IL(T, W, X) hipLaunchKernelGGL((k_interleave<T, W>), g, t, 0, s, l, n, X)
  if (x64) {
    if (width == 1) QMBED_IL(double, 1, x64);
    else if (width == 2) QMBED_IL(double, 2, x64);
    else if (width == 4) QMBED_IL(double, 4, x64);
    else if (width == 8) QMBED_IL(double, 8, x64);
    else QMBED_IL(double, 16, x64);
  } else {
    if (width == 1) QMBED_IL(float, 1, x32);
    else if (width == 2) QMBED_IL(float, 2, x32);
    else if (width == 4) QMBED_IL(float, 4, x32);
    else if (width == 8) QMBED_IL(float, 8, x32);
    else QMBED_IL(float, 16, x32);
  }
#undef QMBED_IL
  QMBED_HIP(hipGetLastError());
}

void scatter_lane(const double* x, size_t n, const Interleave& il, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter_lane, dim3(nblocks(n)), dim3(kT), 0, s, x, n, il);
  QMBED_HIP(hipGetLastError());
}

}  // namespace lk

void deterministic_start_device(double* d_out, size_t n, uint64_t seed, hipStream_t s, bool complex) {
  hipLaunchKernelGGL(lk::k_det_start, dim3(lk::nblocks(n)), dim3(lk::kT), 0, s, d_out, n, seed, complex ? 1 : 0);
  QMBED_HIP(hipGetLastError());
}

}  // namespace qmbed
