// Batched Krylov exponential actions on gfx950: many right-hand sides of one
// Hamiltonian advance in lockstep so every H·v launch serves all columns
// (the orbit walk and index lookup of a row are shared by the columns).
//
//  * evolve_batch: adaptive Lanczos-projection evolution with the Simpson
//    a-posteriori bound (QMBED src/solve/mod.rs:2909-3316, 4141-4170);
//    per-column step control is host logic on m×m projected data, all
//    N-length work is on the device.
//  * ExpmActionPlan: truncated-Taylor action with scaling, trace shift and
//    the ∞-norm early stop (src/solve/mod.rs:59-248); the trace and one-norm
//    come from a row-wise exact-entry kernel (H Hermitian ⇒ ‖A‖₁ = ‖A‖∞).
//
// Layout: a block of `nc` columns is contiguous ([col][row], complex
// interleaved), Krylov basis V is [j][col][row].
#include <algorithm>
#include <cmath>
#include <exception>
#include <thread>

#include "internal.hpp"

namespace qmbed {

std::vector<double> tridiagonal_eigvec(const std::vector<double>& alpha, const std::vector<double>& beta, double lam);

namespace {

constexpr int kT = 256;
constexpr double F64_EPS = 2.220446049250313e-16;

int row_blocks(size_t n) { return int(std::min<size_t>(256, std::max<size_t>(1, (n + kT * 4 - 1) / (kT * 4)))); }

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
__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kT / 64; ++w) r = fmax(r, sh[w]);
  __syncthreads();
  return r;
}

// partial[(col*nb + b)*2 + {0,1}] = Σ conj(x) y over block rows (per column)
__global__ void kb_zdot(const double2* __restrict__ X, const double2* __restrict__ Y, size_t n, double* partial) {
  __shared__ double sh[4];
  const int col = blockIdx.y;
  const double2* x = X + size_t(col) * n;
  const double2* y = Y + size_t(col) * n;
  double sr = 0, si = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 a = x[i], b = y[i];
    sr += a.x * b.x + a.y * b.y;
    si += a.x * b.y - a.y * b.x;
  }
  sr = block_sum(sr, sh);
  si = block_sum(si, sh);
  if (threadIdx.x == 0) {
    partial[(size_t(col) * gridDim.x + blockIdx.x) * 2] = sr;
    partial[(size_t(col) * gridDim.x + blockIdx.x) * 2 + 1] = si;
  }
}

// per-column |x|^2 (MODE 0) or max |x_i| (MODE 1)
template <int MODE>
__global__ void kb_norm(const double2* __restrict__ X, size_t n, double* partial) {
  __shared__ double sh[4];
  const int col = blockIdx.y;
  const double2* x = X + size_t(col) * n;
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 a = x[i];
    double v = a.x * a.x + a.y * a.y;
    s = MODE ? fmax(s, sqrt(v)) : s + v;
  }
  s = MODE ? block_max(s, sh) : block_sum(s, sh);
  if (threadIdx.x == 0) partial[size_t(col) * gridDim.x + blockIdx.x] = s;
}

// out[col*w + k] = Σ_b partial[(col*nb + b)*w + k]  (MODE 1: max)
template <int MODE>
__global__ void kb_reduce(const double* partial, int nb, int w, int ncol, double* out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * w) return;
  int col = t / w, k = t % w;
  double s = 0;
  for (int b = 0; b < nb; ++b) {
    double v = partial[(size_t(col) * nb + b) * w + k];
    s = MODE ? fmax(s, v) : s + v;
  }
  out[t] = s;
}

// out = scale[col] * X (per column real scale; zero when scale is 0)
__global__ void kb_scale(const double2* X, const double* scale, const int* active, size_t n, double2* out) {
  const int col = blockIdx.y;
  if (active && !active[col]) return;
  const double s = scale[col];
  const double2* x = X + size_t(col) * n;
  double2* o = out + size_t(col) * n;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 a = x[i];
    o[i] = make_double2(s * a.x, s * a.y);
  }
}

// Taylor step: applied = f/order · (applied − μ term); term <- applied;
// sum += term  (per column complex factor; inactive columns untouched)
__global__ void kb_taylor(double2* applied, double2* term, double2* sum, const double* fr, const double* fi,
                          double mur, double mui, const int* active, size_t n) {
  const int col = blockIdx.y;
  if (!active[col]) return;
  const double a = fr[col], b = fi[col];
  const size_t off = size_t(col) * n;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 ap = applied[off + i], t = term[off + i];
    double xr = ap.x - (mur * t.x - mui * t.y);
    double xi = ap.y - (mur * t.y + mui * t.x);
    double2 nt = make_double2(a * xr - b * xi, a * xi + b * xr);
    term[off + i] = nt;
    double2 s = sum[off + i];
    sum[off + i] = make_double2(s.x + nt.x, s.y + nt.y);
  }
}

// state = eta · sum  (eta complex, uniform)
__global__ void kb_eta(const double2* sum, double er, double ei, const int* active, size_t n, double2* state) {
  const int col = blockIdx.y;
  if (!active[col]) return;
  const size_t off = size_t(col) * n;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 s = sum[off + i];
    state[off + i] = make_double2(er * s.x - ei * s.y, er * s.y + ei * s.x);
  }
}

struct Batch {
  size_t n;
  int nc;
  hipStream_t s;
  DeviceArray<double> partial, out;
  std::vector<double> host;
  Batch(size_t n_, int nc_, hipStream_t s_) : n(n_), nc(nc_), s(s_) {
    partial.resize(size_t(nc) * 256 * 2 + 64);
    out.resize(size_t(nc) * 2 + 64);
    host.resize(size_t(nc) * 2 + 64);
  }
  dim3 grid() const { return dim3(row_blocks(n), nc); }
  // per-column complex dots conj(X)·Y -> host (re, im) pairs
  void zdot(const cplx* X, const cplx* Y) {
    int nb = row_blocks(n);
    hipLaunchKernelGGL(kb_zdot, grid(), dim3(kT), 0, s, reinterpret_cast<const double2*>(X),
                       reinterpret_cast<const double2*>(Y), n, partial.data());
    hipLaunchKernelGGL(kb_reduce<0>, dim3((nc * 2 + 63) / 64), dim3(64), 0, s, partial.data(), nb, 2, nc, out.data());
    QMBED_HIP(hipMemcpyAsync(host.data(), out.data(), size_t(nc) * 2 * 8, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
  }
  // per-column norms^2 (MODE 0) or ∞-norms (MODE 1) -> host
  void norms(const cplx* X, int mode) {
    int nb = row_blocks(n);
    if (mode == 0)
      hipLaunchKernelGGL(kb_norm<0>, grid(), dim3(kT), 0, s, reinterpret_cast<const double2*>(X), n, partial.data());
    else
      hipLaunchKernelGGL(kb_norm<1>, grid(), dim3(kT), 0, s, reinterpret_cast<const double2*>(X), n, partial.data());
    if (mode == 0)
      hipLaunchKernelGGL(kb_reduce<0>, dim3((nc + 63) / 64), dim3(64), 0, s, partial.data(), nb, 1, nc, out.data());
    else
      hipLaunchKernelGGL(kb_reduce<1>, dim3((nc + 63) / 64), dim3(64), 0, s, partial.data(), nb, 1, nc, out.data());
    QMBED_HIP(hipMemcpyAsync(host.data(), out.data(), size_t(nc) * 8, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
  }
};

// ---------------------------------------------------- projected algebra ----
struct Projection {
  double init_norm = 0;
  int size = 0;
  double residual_beta = 0;
  std::vector<double> evals, evecs;  // evecs row-major size×size
};

std::vector<cplx> projected_coefficients(const Projection& p, double interval) {
  // src/solve/mod.rs:3056-3076
  std::vector<cplx> c(p.size, cplx(0, 0));
  for (int e = 0; e < p.size; ++e) {
    cplx ex = std::exp(cplx(0.0, -interval * p.evals[e]));
    cplx w = p.init_norm * p.evecs[e] * ex;
    for (int b = 0; b < p.size; ++b) c[b] += p.evecs[size_t(b) * p.size + e] * w;
  }
  return c;
}

double residual_amplitude(const Projection& p, double interval) {
  // src/solve/mod.rs:3131-3144
  if (p.size == 0) return 0.0;
  cplx a(0, 0);
  const int last = p.size - 1;
  for (int e = 0; e < p.size; ++e)
    a += p.init_norm * p.evecs[e] * p.evecs[size_t(last) * p.size + e] * std::exp(cplx(0, -interval * p.evals[e]));
  return std::abs(a);
}

double error_bound(const Projection& p, double interval, size_t n) {
  // src/solve/mod.rs:3153-3195
  double dur = std::fabs(interval);
  if (dur <= F64_EPS || p.size == 0 || size_t(p.size) == n || p.residual_beta <= 1.0e-14) return 0.0;
  double mn = INFINITY, mx = -INFINITY;
  for (double v : p.evals) {
    mn = std::min(mn, v);
    mx = std::max(mx, v);
  }
  double span = dur * std::max(mx - mn, 0.0);
  long osc = long(std::min(std::ceil(span / M_PI), 60.0));
  long intervals = std::min<long>(16 + 4 * osc, 256);
  if (intervals % 2) ++intervals;
  double dir = interval > 0 ? 1.0 : (interval < 0 ? -1.0 : 0.0);
  double step = dur / double(intervals);
  double integral = 0;
  for (long i = 0; i <= intervals; ++i) {
    double w = (i == 0 || i == intervals) ? 1.0 : (i % 2 == 0 ? 2.0 : 4.0);
    integral += w * residual_amplitude(p, dir * step * double(i));
  }
  return 1.25 * p.residual_beta * step * integral / 3.0;
}

}  // namespace

void EvolutionOptions::validate() const {
  bool bad_times = times.empty();
  for (size_t i = 0; i < times.size() && !bad_times; ++i)
    if (!std::isfinite(times[i]) || (i && times[i - 1] > times[i])) bad_times = true;
  if (bad_times) throw invalid_options("times must be a nonempty finite nondecreasing grid");
  if (krylov_dimension <= 0 || !std::isfinite(tolerance) || tolerance <= 0.0 || max_substeps <= 0)
    throw invalid_options("evolution controls must be positive");
}

namespace {

// ------------------------------------------- device-resident projections ---
// One projection of all columns runs without a host round trip: α_j, β_j and
// the stop of each column (window end or breakdown β <= 1e-14) are decided on
// the device and kept in a per-column history that the host reads once, at
// the end of the projection.
//
// The columns are interleaved: element (row r, column c) of a block of nc
// columns at r·nc + c, so the H·v's gather of a target row fetches the inputs
// of several columns in one cache line. Element-wise kernels map a block of
// 256 threads onto 4 rows × 64 columns (consecutive threads, consecutive
// columns of one row: coalesced); per-column sums combine the 4 row lanes in
// LDS and leave one partial per row block, partial[c · nb + block].

constexpr int kCols = 64;              // column lanes of an element-wise block
constexpr int kRowLanes = kT / kCols;  // row lanes

int il_row_blocks(size_t n) {
  return int(std::min<size_t>(1024, std::max<size_t>(1, (n + size_t(kRowLanes) * 16 - 1) / (size_t(kRowLanes) * 16))));
}
dim3 il_grid(size_t n, int cols) { return dim3(il_row_blocks(n), (cols + kCols - 1) / kCols); }

// per-column sum of the 4 row lanes into partial[c · nb + block]
__device__ __forceinline__ void il_partial(double v, int nc, double* sh, double* partial, int stride, int off) {
  sh[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x < kCols) {
    const int c = blockIdx.y * kCols + threadIdx.x;
    double t = sh[threadIdx.x];
    for (int q = 1; q < kRowLanes; ++q) t += sh[q * kCols + threadIdx.x];
    if (c < nc) partial[(size_t(c) * gridDim.x + blockIdx.x) * stride + off] = t;
  }
  __syncthreads();
}

// Σ_r |X[r, c]|² per column (active == nullptr: every column)
__global__ __launch_bounds__(kT) void il_norm(const double2* __restrict__ X, size_t n, int nc, const int* active,
                                              double* partial) {
  __shared__ double sh[kT];
  const int c = blockIdx.y * kCols + (threadIdx.x % kCols);
  double acc = 0;
  if (c < nc && (!active || active[c]))
    for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes) {
      const double2 v = X[r * nc + c];
      acc += v.x * v.x + v.y * v.y;
    }
  il_partial(acc, nc, sh, partial, 1, 0);
}

// Σ_r conj(X[r, c]) Y[r, c] per column: (re, im) pairs in the partials
__global__ __launch_bounds__(kT) void il_zdot(const double2* __restrict__ X, const double2* __restrict__ Y, size_t n,
                                              int nc, const int* active, double* partial) {
  __shared__ double sh[kT];
  const int c = blockIdx.y * kCols + (threadIdx.x % kCols);
  double sr = 0, si = 0;
  if (c < nc && (!active || active[c]))
    for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes) {
      const double2 a = X[r * nc + c], b = Y[r * nc + c];
      sr += a.x * b.x + a.y * b.y;
      si += a.x * b.y - a.y * b.x;
    }
  il_partial(sr, nc, sh, partial, 2, 0);
  il_partial(si, nc, sh, partial, 2, 1);
}

// out[r, c] = scale[c] · X[r, c] for the active columns (active == nullptr:
// every column)
__global__ __launch_bounds__(kT) void il_scale(const double2* X, const double* scale, const int* active, size_t n,
                                               int nc, double2* out) {
  const int c = blockIdx.y * kCols + (threadIdx.x % kCols);
  if (c >= nc || (active && !active[c])) return;
  const double f = scale[c];
  for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes) {
    const double2 v = X[r * nc + c];
    out[r * nc + c] = make_double2(f * v.x, f * v.y);
  }
}

// W -= α_j V_j + β_{j−1} V_{j−1} with the coefficients from the history, and
// the partials of ‖W‖² of the result
__global__ __launch_bounds__(kT) void il_three(double2* W, const double2* __restrict__ V, const double2* __restrict__ U,
                                               const double* are, const double* bh, const int* running, int m, int j,
                                               size_t n, int nc, double* partial) {
  __shared__ double sh[kT];
  const int c = blockIdx.y * kCols + (threadIdx.x % kCols);
  double acc = 0;
  if (c < nc && running[c]) {
    const double a = are[size_t(c) * m + j];
    const double b = j > 0 ? bh[size_t(c) * m + j - 1] : 0.0;
    for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes) {
      const size_t i = r * nc + c;
      double2 x = W[i];
      const double2 y = V[i];
      x.x -= a * y.x;
      x.y -= a * y.y;
      if (U) {
        const double2 z = U[i];
        x.x -= b * z.x;
        x.y -= b * z.y;
      }
      W[i] = x;
      acc += x.x * x.x + x.y * x.y;
    }
  }
  il_partial(acc, nc, sh, partial, 1, 0);
}

// [col][row] <-> interleaved (rows r of columns c0.., count columns)
__global__ __launch_bounds__(kT) void il_from_cols(const double2* __restrict__ cols, size_t n, int nc, int c0,
                                                   int count, double2* __restrict__ il) {
  const int v = blockIdx.y * kCols + (threadIdx.x % kCols);
  if (v >= count) return;
  for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes)
    il[r * nc + c0 + v] = cols[size_t(v) * n + r];
}
__global__ __launch_bounds__(kT) void il_to_cols(const double2* __restrict__ il, size_t n, int nc, int c0, int count,
                                                 double2* __restrict__ cols) {
  const int v = blockIdx.y * kCols + (threadIdx.x % kCols);
  if (v >= count) return;
  for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes)
    cols[size_t(v) * n + r] = il[r * nc + c0 + v];
}

// Per-column reductions of the row-block partials run one wave per column:
// lane-strided sums, then a shuffle tree (a fixed order).
__device__ __forceinline__ double wave_total(const double* p, int nb, int stride, int lane) {
  double s = 0;
  for (int b = lane; b < nb; b += 64) s += p[size_t(b) * stride];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  return s;
}

// ‖cur_c‖ from the norm partials: init_norm, running = act && norm > eps,
// inv = 1/norm (0 when not running), size = 0
__global__ void kb_proj_init(const double* partial, int nb, int nc, const int* act, double* init_norm, int* running,
                             double* inv, int* size) {
  const int c = blockIdx.x * (kT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= nc) return;
  const double s = wave_total(partial + size_t(c) * nb, nb, 1, lane);
  if (lane) return;
  const double nrm = sqrt(s);
  init_norm[c] = nrm;
  const bool run = act[c] && nrm > F64_EPS;
  running[c] = run ? 1 : 0;
  inv[c] = run ? 1.0 / nrm : 0.0;
  size[c] = 0;
}

// α_j (real and imaginary part) of the running columns into the history
__global__ void kb_alpha_hist(const double* partial, int nb, int nc, const int* running, int m, int j, double* are,
                              double* aim) {
  const int c = blockIdx.x * (kT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= nc || !running[c]) return;
  const double sr = wave_total(partial + size_t(c) * nb * 2, nb, 2, lane);
  const double si = wave_total(partial + size_t(c) * nb * 2 + 1, nb, 2, lane);
  if (lane) return;
  are[size_t(c) * m + j] = sr;
  aim[size_t(c) * m + j] = si;
}

// β_j = ‖w‖ into the history; the column stops at the window's end or at a
// breakdown (src/solve/mod.rs:2909-3054), with size = j + 1
__global__ void kb_beta_hist(const double* partial, int nb, int nc, int* running, int m, int j, double* bh,
                             int* size, double* inv) {
  const int c = blockIdx.x * (kT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= nc || !running[c]) return;
  const double s = wave_total(partial + size_t(c) * nb, nb, 1, lane);
  if (lane) return;
  const double beta = sqrt(s);
  bh[size_t(c) * m + j] = beta;
  if (j + 1 == m || beta <= 1.0e-14) {
    running[c] = 0;
    size[c] = j + 1;
    inv[c] = 0.0;
  } else {
    inv[c] = 1.0 / beta;
  }
}

// One output or state update of column `col`: dst[r, col] = Σ_j coef[j]
// V[j][r, col] over `size` Krylov blocks, renormalized to the column's
// initial norm (a separate pass once the norms are known). Copy actions
// (size < 0) copy src[r, col] instead.
struct Act {
  double2* dst;        // interleaved block of nc columns
  const double2* src;  // copy actions: interleaved block to copy from
  int col;
  int size;
  int coef;  // offset (complex entries) into the coefficient array
  int pad;
};

// Thread lanes run over actions (consecutive actions: consecutive columns of
// one row, in practice), with the per-action partials of ‖dst‖²
__global__ __launch_bounds__(kT) void il_combine_acts(const Act* __restrict__ acts, int nacts,
                                                      const double2* __restrict__ V, size_t ldj, int nc,
                                                      const double* __restrict__ coefs, size_t n, double* partial) {
  __shared__ double sh[kT];
  const int ai = blockIdx.y * kCols + (threadIdx.x % kCols);
  double acc = 0;
  if (ai < nacts) {
    const Act a = acts[ai];
    const double* cf = coefs + 2 * size_t(a.coef);
    for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes) {
      const size_t i = r * nc + a.col;
      double ar = 0, aim = 0;
      for (int j = 0; j < a.size; ++j) {
        const double2 v = V[size_t(j) * ldj + i];
        ar += v.x * cf[2 * j] - v.y * cf[2 * j + 1];
        aim += v.x * cf[2 * j + 1] + v.y * cf[2 * j];
      }
      a.dst[i] = make_double2(ar, aim);
      acc += ar * ar + aim * aim;
    }
  }
  il_partial(acc, nacts, sh, partial, 1, 0);
}

// per-action factor init_norm[col] / ‖dst‖ (1 when that norm is not finite
// and above eps: src/solve/mod.rs:3104-3112), one wave per action
__global__ void kb_act_scale(const Act* __restrict__ acts, int nacts, const double* partial, int nb,
                             const double* init_norm, double* scale) {
  const int ai = blockIdx.x * (kT / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (ai >= nacts) return;
  const double t = wave_total(partial + size_t(ai) * nb, nb, 1, lane);
  if (lane) return;
  const double on = sqrt(t);
  scale[ai] = (on > F64_EPS && isfinite(on)) ? init_norm[acts[ai].col] / on : 1.0;
}

__global__ __launch_bounds__(kT) void il_scale_acts(const Act* __restrict__ acts, int nacts, const double* scale,
                                                    size_t n, int nc) {
  const int ai = blockIdx.y * kCols + (threadIdx.x % kCols);
  if (ai >= nacts) return;
  const Act a = acts[ai];
  const double f = scale[ai];
  if (f == 1.0) return;
  for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes) {
    const size_t i = r * nc + a.col;
    const double2 v = a.dst[i];
    a.dst[i] = make_double2(f * v.x, f * v.y);
  }
}

__global__ __launch_bounds__(kT) void il_copy_acts(const Act* __restrict__ acts, int nacts, size_t n, int nc) {
  const int ai = blockIdx.y * kCols + (threadIdx.x % kCols);
  if (ai >= nacts) return;
  const Act a = acts[ai];
  for (size_t r = blockIdx.x * size_t(kRowLanes) + threadIdx.x / kCols; r < n; r += size_t(gridDim.x) * kRowLanes)
    a.dst[r * nc + a.col] = a.src[r * nc + a.col];
}

// per-column Σ |Im x| partials of [col][row] input (a column with none is a
// real state)
__global__ void kb_imag_abs(const double2* __restrict__ X, size_t n, double* partial) {
  __shared__ double sh[4];
  const double2* x = X + size_t(blockIdx.y) * n;
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) s += fabs(x[i].y);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// host threads for the per-column step control: the hardware threads, at
// most 16
int control_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return int(std::max(1u, std::min(16u, h ? h : 1u)));
}

}  // namespace

void evolve_batch_device(const Operator& op, const cplx* d_cols, int nc, const EvolutionOptions& opts,
                         EvolutionDiagnostics* diag, cplx* d_out, hipStream_t s) {
  opts.validate();
  const size_t n = op.dim();
  if (!opts.hamiltonian) throw unsupported_backend("direct-generator evolution is outside the MI355X path");
  const int T = int(opts.times.size());
  EvolutionDiagnostics dg;
  if (nc <= 0 || n == 0) {
    if (diag) *diag = dg;
    return;
  }
  const int m = int(std::max<long>(std::min<long>(opts.krylov_dimension, long(n)), 1));
  if (m > 128) throw unsupported_backend("Krylov dimension above 128 exceeds the batched projection kernels");
  const double local_tol = opts.tolerance / double(std::max(T, 1));
  const int nb = il_row_blocks(n);
  const dim3 grid = il_grid(n, nc);
  // the interleaved H·v (compiled stream); otherwise each column block is
  // moved to column-major and back around the column-major multi-vector H·v
  const bool il_action = prepare_interleaved_action(op, s);
  const size_t blk = n * size_t(nc);  // complex elements per interleaved block

  DeviceArray<double> cur(2 * blk), W(2 * blk), V(2 * blk * size_t(m)), outs(2 * blk * size_t(T));
  DeviceArray<double> tmp_x, tmp_y;
  if (!il_action) {
    tmp_x.resize(2 * blk);
    tmp_y.resize(2 * blk);
  }
  const int max_acts = nc * (T + 1);
  DeviceArray<double> partial(size_t(std::max(max_acts, 2 * nc)) * size_t(std::max(nb, 256)) + 64),
      are(size_t(nc) * m), aim(size_t(nc) * m), bh(size_t(nc) * m), init_norm(nc), inv(nc), ascale(max_acts),
      coefs(2 * size_t(nc) * (T + 1) * m + 2);
  DeviceArray<int> act_d(nc), running(nc), size_d(nc);
  DeviceArray<uint8_t> acts_d(sizeof(Act) * (size_t(max_acts) + size_t(nc)) + sizeof(Act));
  PinnedArray<double> hist;  // are | aim | bh | init_norm | size (as int pairs)
  hist.ensure(3 * size_t(nc) * m + 2 * size_t(nc) + 8);
  PinnedArray<int> act_h;
  act_h.ensure(nc);
  PinnedArray<Act> acts_h;
  acts_h.ensure(size_t(max_acts) + size_t(nc) + 1);
  PinnedArray<Act> emit_h;
  emit_h.ensure(size_t(nc) * T + 1);
  DeviceArray<uint8_t> emit_d(sizeof(Act) * (size_t(nc) * T + 1));
  PinnedArray<double> coef_h;
  coef_h.ensure(2 * size_t(nc) * (T + 1) * m + 2);

  double2* curp = reinterpret_cast<double2*>(cur.data());
  double2* Wp = reinterpret_cast<double2*>(W.data());
  double2* Vc = reinterpret_cast<double2*>(V.data());
  auto out_blk = [&](int q) { return reinterpret_cast<double2*>(outs.data()) + blk * size_t(q); };

  // real initial states (for the real-projection counters and the α check)
  std::vector<int> real_state(nc, 1);
  {
    const int nbc = row_blocks(n);
    hipLaunchKernelGGL(kb_imag_abs, dim3(nbc, nc), dim3(kT), 0, s, reinterpret_cast<const double2*>(d_cols), n,
                       partial.data());
    std::vector<double> h(size_t(nc) * nbc);
    QMBED_HIP(hipMemcpyAsync(h.data(), partial.data(), h.size() * 8, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < nc; ++c)
      for (int b = 0; b < nbc; ++b)
        if (h[size_t(c) * nbc + b] != 0.0) real_state[c] = 0;
  }
  hipLaunchKernelGGL(il_from_cols, grid, dim3(kT), 0, s, reinterpret_cast<const double2*>(d_cols), n, nc, 0, nc,
                     curp);
  std::vector<double> t(nc, 0.0);
  std::vector<int> oi(nc, 0), substeps(nc, 0), done(nc, 0);
  const size_t ldj = blk;  // complex elements per Krylov block
  for (;;) {
    // emit outputs whose time is reached; check budgets
    int nact = 0, nemit = 0;
    for (int c = 0; c < nc; ++c) {
      act_h.p[c] = 0;
      if (done[c]) continue;
      while (oi[c] < T) {
        const double target = opts.times[oi[c]];
        if (std::fabs(target - t[c]) <= 16.0 * F64_EPS * std::max(std::fabs(target), 1.0)) {
          Act e{};
          e.dst = out_blk(oi[c]);
          e.src = curp;
          e.col = c;
          emit_h.p[nemit++] = e;
          t[c] = target;
          ++oi[c];
        } else {
          break;
        }
      }
      if (oi[c] >= T) {
        done[c] = 1;
        continue;
      }
      if (substeps[c] >= opts.max_substeps) throw non_convergence(substeps[c], std::fabs(opts.times[oi[c]] - t[c]));
      act_h.p[c] = 1;
      ++nact;
    }
    if (nemit > 0) {
      QMBED_HIP(hipMemcpyAsync(emit_d.data(), emit_h.p, sizeof(Act) * size_t(nemit), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(il_copy_acts, il_grid(n, nemit), dim3(kT), 0, s, reinterpret_cast<const Act*>(emit_d.data()),
                         nemit, n, nc);
      // the emission list is rewritten only after the next synchronization
    }
    if (nact == 0) break;
    // ---- batched Lanczos projection (no reorthogonalization) ----
    QMBED_HIP(hipMemcpyAsync(act_d.data(), act_h.p, size_t(nc) * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(il_norm, grid, dim3(kT), 0, s, curp, n, nc, nullptr, partial.data());
    hipLaunchKernelGGL(kb_proj_init, dim3((nc + 3) / 4), dim3(kT), 0, s, partial.data(), nb, nc, act_d.data(),
                       init_norm.data(), running.data(), inv.data(), size_d.data());
    hipLaunchKernelGGL(il_scale, grid, dim3(kT), 0, s, curp, inv.data(), nullptr, n, nc, Vc);
    // runs of consecutive active columns share one multi-column H·v
    std::vector<std::pair<int, int>> runs;
    for (int c = 0; c < nc;) {
      if (!act_h.p[c]) {
        ++c;
        continue;
      }
      int e = c;
      while (e < nc && act_h.p[e]) ++e;
      runs.push_back({c, e - c});
      c = e;
    }
    for (int j = 0; j < m; ++j) {
      double2* Vj = Vc + size_t(j) * ldj;
      for (auto& r : runs) {
        if (il_action) {
          op.apply_batch_interleaved(reinterpret_cast<const cplx*>(Vj + r.first), reinterpret_cast<cplx*>(Wp + r.first),
                                     size_t(nc), r.second, s);
        } else {
          const dim3 g2 = il_grid(n, r.second);
          double2* tx = reinterpret_cast<double2*>(tmp_x.data());
          double2* ty = reinterpret_cast<double2*>(tmp_y.data());
          hipLaunchKernelGGL(il_to_cols, g2, dim3(kT), 0, s, Vj, n, nc, r.first, r.second, tx);
          op.apply_batch(reinterpret_cast<const cplx*>(tx), reinterpret_cast<cplx*>(ty), r.second, s);
          hipLaunchKernelGGL(il_from_cols, g2, dim3(kT), 0, s, ty, n, nc, r.first, r.second, Wp);
        }
      }
      hipLaunchKernelGGL(il_zdot, grid, dim3(kT), 0, s, Vj, Wp, n, nc, running.data(), partial.data());
      hipLaunchKernelGGL(kb_alpha_hist, dim3((nc + 3) / 4), dim3(kT), 0, s, partial.data(), nb, nc, running.data(),
                         m, j, are.data(), aim.data());
      hipLaunchKernelGGL(il_three, grid, dim3(kT), 0, s, Wp, Vj, j > 0 ? Vj - ldj : nullptr, are.data(), bh.data(),
                         running.data(), m, j, n, nc, partial.data());
      hipLaunchKernelGGL(kb_beta_hist, dim3((nc + 3) / 4), dim3(kT), 0, s, partial.data(), nb, nc, running.data(),
                         m, j, bh.data(), size_d.data(), inv.data());
      if (j + 1 < m)
        hipLaunchKernelGGL(il_scale, grid, dim3(kT), 0, s, Wp, inv.data(), running.data(), n, nc, Vj + ldj);
    }
    QMBED_HIP(hipGetLastError());
    double* hp = hist.p;
    const size_t cm = size_t(nc) * m;
    QMBED_HIP(hipMemcpyAsync(hp, are.data(), cm * 8, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipMemcpyAsync(hp + cm, aim.data(), cm * 8, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipMemcpyAsync(hp + 2 * cm, bh.data(), cm * 8, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipMemcpyAsync(hp + 3 * cm, init_norm.data(), size_t(nc) * 8, hipMemcpyDeviceToHost, s));
    int* sizes = reinterpret_cast<int*>(hp + 3 * cm + nc);
    QMBED_HIP(hipMemcpyAsync(sizes, size_d.data(), size_t(nc) * 4, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
    // ---- per-column step control (host, m×m) ----
    // Columns are independent: their plans are made on several host threads
    // and assembled in column order (the same actions and diagnostics for
    // any thread count).
    struct Plan {
      Projection p;
      std::vector<std::pair<double2*, double>> outs;  // (destination block, interval)
      const double2* copy_src = nullptr;              // state <- last output
      double t = 0;
      int oi = 0;
      bool stepped = false;
      bool nonherm = false;
      EvolutionDiagnostics dg;
    };
    std::vector<int> cols_act;
    for (int c = 0; c < nc; ++c)
      if (act_h.p[c]) cols_act.push_back(c);
    std::vector<Plan> plans(cols_act.size());
    auto plan_column = [&](size_t idx) {
      const int c = cols_act[idx];
      Plan& pl = plans[idx];
      Projection& p = pl.p;
      EvolutionDiagnostics& d = pl.dg;
      pl.t = t[c];
      pl.oi = oi[c];
      p.init_norm = hp[3 * cm + c];
      p.size = p.init_norm <= F64_EPS ? 0 : sizes[c];
      const double* ac = hp + size_t(c) * m;
      const double* ic = hp + cm + size_t(c) * m;
      const double* bc = hp + 2 * cm + size_t(c) * m;
      if (!(op.is_real() && real_state[c]))
        for (int j = 0; j < p.size; ++j)
          if (std::fabs(ic[j]) > 1.0e-10) {
            pl.nonherm = true;
            return;
          }
      if (p.size > 0) {
        std::vector<double> aa(ac, ac + p.size), bb(bc, bc + (p.size - 1));
        p.residual_beta = bc[p.size - 1];
        tridiagonal_eigen(aa, bb, p.evals, &p.evecs);
      }
      ++d.lanczos_projections;
      d.matrix_vector_products += p.size;
      if (op.is_real() && real_state[c]) {
        ++d.real_lanczos_projections;
        d.real_matrix_vector_products += p.size;
      }
      const double thr = local_tol * std::max(p.init_norm, 1.0);
      int accepted = 0;
      double last_time = pl.t;
      for (int q = pl.oi; q < T; ++q) {
        const double iv = opts.times[q] - pl.t;
        const double err = error_bound(p, iv, n);
        d.maximum_estimated_error = std::max(d.maximum_estimated_error, err);
        if (err > thr) {
          ++d.rejected_trial_intervals;
          break;
        }
        pl.outs.push_back({out_blk(q), iv});
        ++accepted;
        last_time = opts.times[q];
      }
      pl.stepped = true;
      ++d.accepted_substeps;
      if (accepted) {
        pl.copy_src = out_blk(pl.oi + accepted - 1);
        pl.t = last_time;
        pl.oi += accepted;
        return;
      }
      const double target = opts.times[pl.oi];
      const double tiv = target - pl.t;
      const double dir = tiv > 0 ? 1.0 : -1.0;
      const double min_iv = 16.0 * F64_EPS * std::max(std::fabs(target), 1.0);
      double rej = std::fabs(tiv), acc = rej;
      for (;;) {
        const double err = error_bound(p, dir * acc, n);
        d.maximum_estimated_error = std::max(d.maximum_estimated_error, err);
        if (err <= thr || acc <= min_iv) break;
        ++d.rejected_trial_intervals;
        rej = acc;
        acc *= 0.5;
      }
      if (acc > min_iv && rej > acc) {
        for (int r = 0; r < 8; ++r) {
          const double trial = 0.5 * (acc + rej);
          const double err = error_bound(p, dir * trial, n);
          d.maximum_estimated_error = std::max(d.maximum_estimated_error, err);
          if (err <= thr) acc = trial;
          else {
            ++d.rejected_trial_intervals;
            rej = trial;
          }
        }
      }
      const double iv = dir * acc;
      pl.t += iv;
      pl.outs.push_back({curp, iv});
    };
    {
      const size_t np = plans.size();
      const size_t nth = std::min<size_t>(np, size_t(control_threads()));
      std::vector<std::thread> pool;
      std::vector<std::exception_ptr> err(nth);
      auto work = [&](size_t tid) {
        try {
          for (size_t i = tid; i < np; i += nth) plan_column(i);
        } catch (...) {
          err[tid] = std::current_exception();
        }
      };
      for (size_t tid = 1; tid < nth; ++tid) pool.emplace_back(work, tid);
      if (nth > 0) work(0);
      for (auto& th : pool) th.join();
      for (auto& e : err)
        if (e) std::rethrow_exception(e);
    }
    int nacts = 0, ncoef = 0;
    std::vector<Act> copies;
    for (size_t idx = 0; idx < plans.size(); ++idx) {
      const int c = cols_act[idx];
      Plan& pl = plans[idx];
      if (pl.nonherm) throw non_hermitian();
      dg.lanczos_projections += pl.dg.lanczos_projections;
      dg.matrix_vector_products += pl.dg.matrix_vector_products;
      dg.real_lanczos_projections += pl.dg.real_lanczos_projections;
      dg.real_matrix_vector_products += pl.dg.real_matrix_vector_products;
      dg.accepted_substeps += pl.dg.accepted_substeps;
      dg.rejected_trial_intervals += pl.dg.rejected_trial_intervals;
      dg.maximum_estimated_error = std::max(dg.maximum_estimated_error, pl.dg.maximum_estimated_error);
      for (auto& o : pl.outs) {
        Act a{};
        a.dst = o.first;
        a.col = c;
        a.size = pl.p.size;
        a.coef = ncoef;
        if (pl.p.size > 0) {
          const auto coef = projected_coefficients(pl.p, o.second);
          for (int jj = 0; jj < pl.p.size; ++jj) {
            coef_h.p[2 * size_t(ncoef + jj)] = coef[jj].real();
            coef_h.p[2 * size_t(ncoef + jj) + 1] = coef[jj].imag();
          }
          ncoef += pl.p.size;
        }
        acts_h.p[nacts++] = a;
      }
      if (pl.copy_src) {
        Act cp{};
        cp.src = pl.copy_src;
        cp.dst = curp;
        cp.col = c;
        copies.push_back(cp);
      }
      t[c] = pl.t;
      oi[c] = pl.oi;
      if (pl.stepped) {
        ++substeps[c];
        real_state[c] = 0;
      }
    }
    if (nacts > 0) {
      const int ncopy = int(copies.size());
      for (int i = 0; i < ncopy; ++i) acts_h.p[nacts + i] = copies[size_t(i)];
      QMBED_HIP(hipMemcpyAsync(acts_d.data(), acts_h.p, sizeof(Act) * size_t(nacts + ncopy), hipMemcpyHostToDevice,
                               s));
      if (ncoef > 0)
        QMBED_HIP(hipMemcpyAsync(coefs.data(), coef_h.p, 16 * size_t(ncoef), hipMemcpyHostToDevice, s));
      const Act* ad = reinterpret_cast<const Act*>(acts_d.data());
      const dim3 ga = il_grid(n, nacts);
      hipLaunchKernelGGL(il_combine_acts, ga, dim3(kT), 0, s, ad, nacts, Vc, ldj, nc, coefs.data(), n,
                         partial.data());
      hipLaunchKernelGGL(kb_act_scale, dim3((nacts + 3) / 4), dim3(kT), 0, s, ad, nacts, partial.data(), nb,
                         init_norm.data(), ascale.data());
      hipLaunchKernelGGL(il_scale_acts, ga, dim3(kT), 0, s, ad, nacts, ascale.data(), n, nc);
      if (ncopy > 0) hipLaunchKernelGGL(il_copy_acts, il_grid(n, ncopy), dim3(kT), 0, s, ad + nacts, ncopy, n, nc);
      QMBED_HIP(hipGetLastError());
    }
  }
  // outputs to the caller's [time][column][row] layout
  for (int q = 0; q < T; ++q)
    hipLaunchKernelGGL(il_to_cols, grid, dim3(kT), 0, s, out_blk(q), n, nc, 0, nc,
                       reinterpret_cast<double2*>(d_out) + blk * size_t(q));
  QMBED_HIP(hipGetLastError());
  QMBED_HIP(hipStreamSynchronize(s));
  if (diag) *diag = dg;
}

std::vector<std::vector<std::vector<cplx>>> evolve_batch(const Operator& op, const std::vector<std::vector<cplx>>& cols,
                                                         const EvolutionOptions& opts, EvolutionDiagnostics* diag,
                                                         hipStream_t s) {
  opts.validate();
  const size_t n = op.dim();
  for (auto& c : cols)
    if (c.size() != n) throw dimension_mismatch("evolution operator and initial state do not match");
  if (!opts.hamiltonian) throw unsupported_backend("direct-generator evolution is outside the MI355X path");
  const int nc = int(cols.size());
  const int T = int(opts.times.size());
  std::vector<std::vector<std::vector<cplx>>> out(T, std::vector<std::vector<cplx>>(nc));
  if (nc == 0) {
    if (diag) *diag = EvolutionDiagnostics();
    return out;
  }
  DeviceArray<double> in(2 * n * nc), res(2 * n * size_t(nc) * T);
  for (int c = 0; c < nc; ++c)
    QMBED_HIP(hipMemcpyAsync(in.data() + 2 * n * c, cols[c].data(), n * 16, hipMemcpyHostToDevice, s));
  evolve_batch_device(op, reinterpret_cast<const cplx*>(in.data()), nc, opts, diag, reinterpret_cast<cplx*>(res.data()),
                      s);
  for (int q = 0; q < T; ++q)
    for (int c = 0; c < nc; ++c) {
      out[q][c].resize(n);
      QMBED_HIP(hipMemcpyAsync(out[q][c].data(), res.data() + 2 * n * (size_t(q) * nc + c), n * 16,
                               hipMemcpyDeviceToHost, s));
    }
  QMBED_HIP(hipStreamSynchronize(s));
  return out;
}

// ------------------------------------------------------- Taylor plan --------
// Al-Mohy & Higham (2011), Table 3.1: θ_m for unit roundoff 2^-53, as
// tabulated by the reference (src/solve/mod.rs:14-50).
static const std::pair<int, double> kTheta[] = {
    {1, 2.29e-16}, {2, 2.58e-8}, {3, 1.39e-5}, {4, 3.40e-4}, {5, 2.40e-3}, {6, 9.07e-3}, {7, 2.38e-2},
    {8, 5.00e-2},  {9, 8.96e-2}, {10, 1.44e-1}, {11, 2.14e-1}, {12, 3.00e-1}, {13, 4.00e-1}, {14, 5.14e-1},
    {15, 6.41e-1}, {16, 7.81e-1}, {17, 9.31e-1}, {18, 1.09}, {19, 1.26}, {20, 1.44}, {21, 1.62},
    {22, 1.82},    {23, 2.01},   {24, 2.22},   {25, 2.43},   {26, 2.64},   {27, 2.86},   {28, 3.08},
    {29, 3.31},    {30, 3.54},   {35, 4.7},    {40, 6.0},    {45, 7.2},    {50, 8.5},    {55, 9.9}};

ExpmActionPlan::ExpmActionPlan(std::shared_ptr<const Operator> op, cplx coefficient, int max_degree, double tolerance,
                               long max_substeps, hipStream_t s)
    : op_(std::move(op)), coefficient_(coefficient), tolerance_(tolerance) {
  if (!std::isfinite(coefficient.real()) || !std::isfinite(coefficient.imag()) || max_degree <= 0 ||
      !std::isfinite(tolerance) || tolerance <= 0.0 || max_substeps <= 0)
    throw invalid_options("invalid exponential coefficient or numerical controls");
  auto tn = op_->shifted_trace_and_one_norm(s);
  shift_ = tn.first;
  one_norm_ = tn.second;
  const double scaled = std::abs(coefficient) * one_norm_;
  degree_ = 0;
  scaling_ = 1;
  if (scaled != 0.0) {
    bool found = false;
    long best = 0;
    for (auto& th : kTheta) {
      if (th.first > max_degree) continue;
      long sc = long(std::max(std::ceil(scaled / th.second), 1.0));
      long cost = long(th.first) * sc;
      if (!found || cost < best) {
        found = true;
        best = cost;
        degree_ = th.first;
        scaling_ = sc;
      }
    }
    if (!found) throw invalid_options("Taylor degree is below the minimum supported degree");
  }
  if (scaling_ > max_substeps) throw non_convergence(max_substeps, scaled);
}

void ExpmActionPlan::apply_device(cplx* d_vec, int nvec, hipStream_t s) const {
  const size_t n = op_->dim();
  if (nvec == 0 || n == 0) return;
  Batch B(n, nvec, s);
  B.norms(d_vec, 1);
  std::vector<int> active(nvec, 1);
  int nact = 0;
  for (int c = 0; c < nvec; ++c) {
    if (std::abs(coefficient_) <= F64_EPS || B.host[c] == 0.0) active[c] = 0;
    nact += active[c];
  }
  if (nact == 0) return;
  const cplx factor = coefficient_ / double(scaling_);
  const cplx eta = std::exp(factor * shift_);
  if (!std::isfinite(eta.real()) || !std::isfinite(eta.imag()))
    throw unsupported_backend("exponential action overflowed its scalar trace shift");
  DeviceArray<double> term(2 * n * nvec), sum(2 * n * nvec), applied(2 * n * nvec), fr(nvec), fi(nvec);
  DeviceArray<int> dact(nvec);
  double2* st = reinterpret_cast<double2*>(d_vec);
  for (long sidx = 0; sidx < scaling_; ++sidx) {
    QMBED_HIP(hipMemcpyAsync(term.data(), d_vec, n * nvec * 16, hipMemcpyDeviceToDevice, s));
    QMBED_HIP(hipMemcpyAsync(sum.data(), d_vec, n * nvec * 16, hipMemcpyDeviceToDevice, s));
    std::vector<int> run = active;
    B.norms(reinterpret_cast<const cplx*>(term.data()), 1);
    std::vector<double> prev(B.host.begin(), B.host.begin() + nvec);
    for (int order = 1; order <= degree_; ++order) {
      int live = 0;
      for (int c = 0; c < nvec; ++c) live += run[c];
      if (!live) break;
      op_->apply_batch(reinterpret_cast<const cplx*>(term.data()), reinterpret_cast<cplx*>(applied.data()), nvec, s);
      cplx sc = factor / double(order);
      std::vector<double> hr(nvec, sc.real()), hi(nvec, sc.imag());
      QMBED_HIP(hipMemcpyAsync(fr.data(), hr.data(), nvec * 8, hipMemcpyHostToDevice, s));
      QMBED_HIP(hipMemcpyAsync(fi.data(), hi.data(), nvec * 8, hipMemcpyHostToDevice, s));
      QMBED_HIP(hipMemcpyAsync(dact.data(), run.data(), nvec * 4, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(kb_taylor, B.grid(), dim3(kT), 0, s, reinterpret_cast<double2*>(applied.data()),
                         reinterpret_cast<double2*>(term.data()), reinterpret_cast<double2*>(sum.data()), fr.data(),
                         fi.data(), shift_.real(), shift_.imag(), dact.data(), n);
      B.norms(reinterpret_cast<const cplx*>(term.data()), 1);
      std::vector<double> tn(B.host.begin(), B.host.begin() + nvec);
      B.norms(reinterpret_cast<const cplx*>(sum.data()), 1);
      for (int c = 0; c < nvec; ++c) {
        if (!run[c]) continue;
        if (prev[c] + tn[c] <= tolerance_ * std::max(B.host[c], 2.2250738585072014e-308)) run[c] = 0;
        else prev[c] = tn[c];
      }
    }
    QMBED_HIP(hipMemcpyAsync(dact.data(), active.data(), nvec * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(kb_eta, B.grid(), dim3(kT), 0, s, reinterpret_cast<const double2*>(sum.data()), eta.real(),
                       eta.imag(), dact.data(), n, st);
  }
  QMBED_HIP(hipStreamSynchronize(s));
}

std::vector<std::vector<cplx>> ExpmActionPlan::apply(const std::vector<std::vector<cplx>>& vectors,
                                                     hipStream_t s) const {
  const size_t n = op_->dim();
  const int nv = int(vectors.size());
  std::vector<std::vector<cplx>> out(nv);
  if (nv == 0) return out;
  DeviceArray<double> d(2 * n * nv);
  for (int c = 0; c < nv; ++c) {
    if (vectors[c].size() != n) throw dimension_mismatch("exponential batch column has the wrong length");
    QMBED_HIP(hipMemcpyAsync(d.data() + 2 * n * c, vectors[c].data(), n * 16, hipMemcpyHostToDevice, s));
  }
  apply_device(reinterpret_cast<cplx*>(d.data()), nv, s);
  for (int c = 0; c < nv; ++c) {
    out[c].resize(n);
    QMBED_HIP(hipMemcpyAsync(out[c].data(), d.data() + 2 * n * c, n * 16, hipMemcpyDeviceToHost, s));
  }
  QMBED_HIP(hipStreamSynchronize(s));
  return out;
}

}  // namespace qmbed
