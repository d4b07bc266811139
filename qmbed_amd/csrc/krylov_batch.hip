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

// Per-column reductions of the row-block partials run one wave per column:
// lane-strided sums, then a shuffle tree (a fixed order).
__device__ __forceinline__ double wave_total(const double* p, int nb, int stride, int lane) {
  double s = 0;
  for (int b = lane; b < nb; b += 64) s += p[size_t(b) * stride];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  return s;
}

// ‖cur_c‖ from the kb_norm partials: init_norm, running = act && norm > eps,
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

// W -= α_j V_j + β_{j−1} V_{j−1} with the coefficients from the history
__global__ void kb_three_hist(double2* W, const double2* __restrict__ V, const double2* __restrict__ U,
                              const double* are, const double* bh, const int* running, int m, int j, size_t n) {
  const int col = blockIdx.y;
  if (!running[col]) return;
  const double a = are[size_t(col) * m + j];
  const double b = j > 0 ? bh[size_t(col) * m + j - 1] : 0.0;
  double2* w = W + size_t(col) * n;
  const double2* v = V + size_t(col) * n;
  const double2* u = U ? U + size_t(col) * n : nullptr;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 x = w[i], y = v[i];
    x.x -= a * y.x;
    x.y -= a * y.y;
    if (u) {
      double2 z = u[i];
      x.x -= b * z.x;
      x.y -= b * z.y;
    }
    w[i] = x;
  }
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

// one output or state update: dst = Σ_j coef[j] V[j] over `size` Krylov
// blocks of column `col` (V: that column in block 0), renormalized to the
// column's initial norm; copy actions (coef < 0) copy V to dst
struct Act {
  const double2* V;
  double2* dst;
  int size;
  int coef;  // offset (complex entries) into the coefficient array
  int col;
  int pad;
};

__global__ void kb_combine_acts(const Act* __restrict__ acts, size_t ldj, const double* __restrict__ coefs, size_t n) {
  __shared__ double shc[2 * 128];
  const Act a = acts[blockIdx.y];
  for (int j = threadIdx.x; j < 2 * a.size; j += kT) shc[j] = coefs[2 * size_t(a.coef) + j];
  __syncthreads();
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double ar = 0, ai = 0;
    for (int j = 0; j < a.size; ++j) {
      const double2 v = a.V[size_t(j) * ldj + i];
      ar += v.x * shc[2 * j] - v.y * shc[2 * j + 1];
      ai += v.x * shc[2 * j + 1] + v.y * shc[2 * j];
    }
    a.dst[i] = make_double2(ar, ai);
  }
}

__global__ void kb_norm_acts(const Act* __restrict__ acts, size_t n, double* partial) {
  __shared__ double sh[4];
  const Act a = acts[blockIdx.y];
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double2 v = a.dst[i];
    s += v.x * v.x + v.y * v.y;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// dst *= init_norm / ‖dst‖ when that norm is finite and above eps
// (src/solve/mod.rs:3104-3112); every block re-adds the partials in order
__global__ void kb_renorm_acts(const Act* __restrict__ acts, const double* partial, int nb,
                               const double* init_norm, size_t n) {
  const Act a = acts[blockIdx.y];
  double t = 0;
  for (int b = 0; b < nb; ++b) t += partial[size_t(blockIdx.y) * nb + b];
  const double on = sqrt(t);
  if (!(on > F64_EPS && isfinite(on))) return;
  const double sc = init_norm[a.col] / on;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double2 v = a.dst[i];
    a.dst[i] = make_double2(sc * v.x, sc * v.y);
  }
}

__global__ void kb_copy_acts(const Act* __restrict__ acts, size_t n) {
  const Act a = acts[blockIdx.y];
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) a.dst[i] = a.V[i];
}

// per-column Σ |Im x| partials (a column with none is a real state)
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

template <typename T>
struct Pinned {
  T* p = nullptr;
  size_t n = 0;
  void ensure(size_t want) {
    if (n >= want) return;
    if (p) QMBED_HIP(hipHostFree(p));
    p = nullptr;
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), want * sizeof(T), hipHostMallocDefault));
    n = want;
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

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
  const int nb = row_blocks(n);
  const dim3 grid(nb, nc);

  DeviceArray<double> cur(2 * n * nc), W(2 * n * nc), V(2 * n * size_t(nc) * m);
  DeviceArray<double> partial(size_t(nc) * nb * std::max(2, T + 1) + 64), are(size_t(nc) * m), aim(size_t(nc) * m),
      bh(size_t(nc) * m), init_norm(nc), inv(nc), coefs(2 * size_t(nc) * (T + 1) * m + 2);
  DeviceArray<int> act_d(nc), running(nc), size_d(nc);
  DeviceArray<uint8_t> acts_d(sizeof(Act) * size_t(nc) * (T + 1) + sizeof(Act));
  Pinned<double> hist;  // are | aim | bh | init_norm | size (as int pairs)
  hist.ensure(3 * size_t(nc) * m + 2 * size_t(nc) + 8);
  Pinned<int> act_h;
  act_h.ensure(nc);
  Pinned<Act> acts_h;
  acts_h.ensure(size_t(nc) * (T + 1) + 1);
  Pinned<double> coef_h;
  coef_h.ensure(2 * size_t(nc) * (T + 1) * m + 2);

  QMBED_HIP(hipMemcpyAsync(cur.data(), d_cols, 16 * n * nc, hipMemcpyDeviceToDevice, s));
  // real initial states (for the real-projection counters and the α check)
  std::vector<int> real_state(nc, 1);
  {
    hipLaunchKernelGGL(kb_imag_abs, grid, dim3(kT), 0, s, reinterpret_cast<const double2*>(cur.data()), n,
                       partial.data());
    std::vector<double> h(size_t(nc) * nb);
    QMBED_HIP(hipMemcpyAsync(h.data(), partial.data(), h.size() * 8, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < nc; ++c)
      for (int b = 0; b < nb; ++b)
        if (h[size_t(c) * nb + b] != 0.0) real_state[c] = 0;
  }
  auto out_col = [&](int q, int c) { return reinterpret_cast<double2*>(d_out) + (size_t(q) * nc + c) * n; };
  auto cur_col = [&](int c) { return reinterpret_cast<double2*>(cur.data()) + size_t(c) * n; };
  std::vector<double> t(nc, 0.0);
  std::vector<int> oi(nc, 0), substeps(nc, 0), done(nc, 0);
  double2* Vc = reinterpret_cast<double2*>(V.data());
  const size_t ldj = n * nc;  // complex elements per Krylov block
  for (;;) {
    // emit outputs whose time is reached; check budgets
    int nact = 0;
    for (int c = 0; c < nc; ++c) {
      act_h.p[c] = 0;
      if (done[c]) continue;
      while (oi[c] < T) {
        const double target = opts.times[oi[c]];
        if (std::fabs(target - t[c]) <= 16.0 * F64_EPS * std::max(std::fabs(target), 1.0)) {
          QMBED_HIP(hipMemcpyAsync(out_col(oi[c], c), cur_col(c), n * 16, hipMemcpyDeviceToDevice, s));
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
    if (nact == 0) break;
    // ---- batched Lanczos projection (no reorthogonalization) ----
    QMBED_HIP(hipMemcpyAsync(act_d.data(), act_h.p, size_t(nc) * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(kb_norm<0>, grid, dim3(kT), 0, s, reinterpret_cast<const double2*>(cur.data()), n,
                       partial.data());
    hipLaunchKernelGGL(kb_proj_init, dim3((nc + 3) / 4), dim3(kT), 0, s, partial.data(), nb, nc, act_d.data(),
                       init_norm.data(), running.data(), inv.data(), size_d.data());
    hipLaunchKernelGGL(kb_scale, grid, dim3(kT), 0, s, reinterpret_cast<const double2*>(cur.data()), inv.data(),
                       nullptr, n, Vc);
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
      for (auto& r : runs)
        op.apply_batch(reinterpret_cast<const cplx*>(Vj + size_t(r.first) * n),
                       reinterpret_cast<cplx*>(W.data()) + size_t(r.first) * n, r.second, s);
      hipLaunchKernelGGL(kb_zdot, grid, dim3(kT), 0, s, Vj, reinterpret_cast<const double2*>(W.data()), n,
                         partial.data());
      hipLaunchKernelGGL(kb_alpha_hist, dim3((nc + 3) / 4), dim3(kT), 0, s, partial.data(), nb, nc, running.data(),
                         m, j, are.data(), aim.data());
      hipLaunchKernelGGL(kb_three_hist, grid, dim3(kT), 0, s, reinterpret_cast<double2*>(W.data()), Vj,
                         j > 0 ? Vj - ldj : nullptr, are.data(), bh.data(), running.data(), m, j, n);
      hipLaunchKernelGGL(kb_norm<0>, grid, dim3(kT), 0, s, reinterpret_cast<const double2*>(W.data()), n,
                         partial.data());
      hipLaunchKernelGGL(kb_beta_hist, dim3((nc + 3) / 4), dim3(kT), 0, s, partial.data(), nb, nc, running.data(),
                         m, j, bh.data(), size_d.data(), inv.data());
      if (j + 1 < m)
        hipLaunchKernelGGL(kb_scale, grid, dim3(kT), 0, s, reinterpret_cast<const double2*>(W.data()), inv.data(),
                           running.data(), n, Vj + ldj);
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
      std::vector<std::pair<double2*, double>> outs;  // (destination, interval)
      double2* copy_src = nullptr;                    // state <- last output
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
        pl.outs.push_back({out_col(q, c), iv});
        ++accepted;
        last_time = opts.times[q];
      }
      pl.stepped = true;
      ++d.accepted_substeps;
      if (accepted) {
        pl.copy_src = out_col(pl.oi + accepted - 1, c);
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
      pl.outs.push_back({cur_col(c), iv});
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
        a.V = Vc + size_t(c) * n;
        a.dst = o.first;
        a.size = pl.p.size;
        a.coef = ncoef;
        a.col = c;
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
        cp.V = pl.copy_src;
        cp.dst = cur_col(c);
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
      hipLaunchKernelGGL(kb_combine_acts, dim3(nb, nacts), dim3(kT), 0, s, ad, ldj, coefs.data(), n);
      hipLaunchKernelGGL(kb_norm_acts, dim3(nb, nacts), dim3(kT), 0, s, ad, n, partial.data());
      hipLaunchKernelGGL(kb_renorm_acts, dim3(nb, nacts), dim3(kT), 0, s, ad, partial.data(), nb, init_norm.data(),
                         n);
      if (ncopy > 0) hipLaunchKernelGGL(kb_copy_acts, dim3(nb, ncopy), dim3(kT), 0, s, ad + nacts, n);
      QMBED_HIP(hipGetLastError());
    }
  }
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
