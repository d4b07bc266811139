// Learned Lanczos warm start for parameter scans of an operator family
// H(c) = H_static + Σ_p c_p H_p (no counterpart in the reference, whose scan
// reuse is EigshWorkspace's balanced subspace start, src/solve/mod.rs:477-526
// and :2419-2444).
//
// Training: the training points are solved by plain Lanczos; their ground
// states stay on the device, gauge-fixed (real: sign of Σ_i ψ_i, or of the
// largest entry when that sum vanishes). One GEMM forms the Gram matrix
// G = Ψᵀ Ψ. Pivoted Cholesky of G (host, n × n) selects the snapshot states
// greedily (each pivot is the state farthest from the span of those before)
// and gives G ≈ L Lᵀ; the snapshot basis S = Ψ_sel R⁻¹ (R = L[sel]ᵀ) is
// orthonormal and row i of L holds the coefficients of ψ_i over S, which are
// the MLP's targets. The MLP (mlp.hip) maps the normalized coefficients c to
// those snapshot coefficients and is trained on the device.
//
// Solving a point: MLP(c) → coefficients y, start vector v₀ = S y, Lanczos
// from v₀. A vanishing prediction falls back to the plain deterministic
// start; the Lanczos residual test decides convergence either way.
//
// Snapshot Rayleigh-Ritz start (eigenvector continuation): training also
// forms the projected parts Sᵀ H_part S (ns H·v and one GEMM per part), so Sᵀ H(c) S = Σ_part c_part Sᵀ H_part S costs nothing per
// point; its lowest eigenvector y (ns × ns Jacobi on the host) gives the
// start S y, the best start the snapshot span holds for the Rayleigh
// quotient. The same map can serve as the network's targets.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <future>
#include <limits>

#include "internal.hpp"
#include "lanczos_kernels.hpp"
#include "scan.hpp"

namespace qmbed {

std::vector<double> tridiagonal_eigvec(const std::vector<double>& alpha, const std::vector<double>& beta, double lam);

namespace {

constexpr int kT = 256;

// partial[b] = Σ x over block b's rows
__global__ __launch_bounds__(kT) void k_sum_blocks(const double* __restrict__ x, size_t n, double* partial) {
  __shared__ double sh[kT / 64];
  double s = 0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) s += x[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kT / 64; ++w) t += sh[w];
    partial[blockIdx.x] = t;
  }
}

// per-block (max |x|, signed value at that max) for the fallback gauge
__global__ __launch_bounds__(kT) void k_absmax_blocks(const double* __restrict__ x, size_t n, double* amax,
                                                      double* aval) {
  __shared__ double shm[kT], shv[kT];
  double m = -1.0, v = 0.0;
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double a = fabs(x[i]);
    if (a > m) {
      m = a;
      v = x[i];
    }
  }
  shm[threadIdx.x] = m;
  shv[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = -1.0, bv = 0.0;
    for (int t = 0; t < kT; ++t)
      if (shm[t] > bm) {
        bm = shm[t];
        bv = shv[t];
      }
    amax[blockIdx.x] = bm;
    aval[blockIdx.x] = bv;
  }
}

__global__ __launch_bounds__(kT) void k_scale_copy(const double* __restrict__ x, size_t n, double f,
                                                   double* __restrict__ y) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) y[i] = f * x[i];
}

__global__ __launch_bounds__(kT) void k_scale(double* __restrict__ x, size_t n, double f) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) x[i] = f * x[i];
}

__global__ __launch_bounds__(kT) void k_fill(double* __restrict__ x, size_t n, double v) {
  for (size_t i = blockIdx.x * size_t(kT) + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) x[i] = v;
}

__global__ void k_f2d(const float* __restrict__ a, int n, double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = double(a[i]);
}

int grid_for(size_t n) { return int(std::min<size_t>(1024, std::max<size_t>(1, (n + kT * 4 - 1) / (kT * 4)))); }

// deterministic uniform numbers in [-1, 1) (64-bit LCG, top 53 bits)
struct Lcg {
  uint64_t x;
  explicit Lcg(uint64_t seed) : x(seed * 2 + 1) {}
  double next() {
    x = x * 6364136223846793005ULL + 1442695040888963407ULL;
    return double(x >> 11) / 4503599627370496.0 - 1.0;
  }
};

// Eigenvector of the lowest eigenvalue of a symmetric n × n matrix
// (column-major): Householder reduction to tridiagonal form, the lowest
// eigenvalue by bisection, its tridiagonal eigenvector by inverse iteration,
// and the reflectors applied back (O(n³) with small constants: tens of
// microseconds at n = 32).
std::vector<double> lowest_eigenvector(std::vector<double> a, int n, double* gap = nullptr) {
  auto A = [&](int i, int j) -> double& { return a[size_t(j) * n + i]; };
  std::vector<std::vector<double>> refl;  // v_k for rows k+1.. (unit, or empty)
  std::vector<double> p(static_cast<size_t>(n)), q(static_cast<size_t>(n));
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    std::vector<double> v(static_cast<size_t>(m));
    double xn = 0;
    for (int i = 0; i < m; ++i) {
      v[size_t(i)] = A(k + 1 + i, k);
      xn += v[size_t(i)] * v[size_t(i)];
    }
    xn = std::sqrt(xn);
    const double alpha = v[0] >= 0 ? -xn : xn;
    v[0] -= alpha;
    double vn = 0;
    for (double x : v) vn += x * x;
    vn = std::sqrt(vn);
    if (vn == 0.0 || xn == 0.0) {
      refl.emplace_back();
      continue;
    }
    for (double& x : v) x /= vn;
    // A22 ← H A22 H, H = I − 2 v vᵀ: p = A22 v, q = p − (vᵀp) v,
    // A22 −= 2 (v qᵀ + q vᵀ)
    for (int i = 0; i < m; ++i) {
      double t = 0;
      for (int j = 0; j < m; ++j) t += A(k + 1 + i, k + 1 + j) * v[size_t(j)];
      p[size_t(i)] = t;
    }
    double kk = 0;
    for (int i = 0; i < m; ++i) kk += v[size_t(i)] * p[size_t(i)];
    for (int i = 0; i < m; ++i) q[size_t(i)] = p[size_t(i)] - kk * v[size_t(i)];
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i)
        A(k + 1 + i, k + 1 + j) -= 2.0 * (v[size_t(i)] * q[size_t(j)] + q[size_t(i)] * v[size_t(j)]);
    A(k + 1, k) = alpha;
    A(k, k + 1) = alpha;
    for (int i = 1; i < m; ++i) A(k + 1 + i, k) = A(k, k + 1 + i) = 0.0;
    refl.push_back(std::move(v));
  }
  std::vector<double> d(static_cast<size_t>(n)), e(static_cast<size_t>(std::max(n - 1, 0)));
  for (int i = 0; i < n; ++i) d[size_t(i)] = A(i, i);
  for (int i = 0; i + 1 < n; ++i) e[size_t(i)] = A(i + 1, i);
  const auto low = tridiagonal_eigenvalue_range(d, e, 0, std::min(n, 2));
  const double lam = low[0];
  if (gap) *gap = low.size() > 1 ? low[1] - low[0] : std::numeric_limits<double>::infinity();
  std::vector<double> z = tridiagonal_eigvec(d, e, lam);
  for (int k = int(refl.size()) - 1; k >= 0; --k) {
    const auto& v = refl[size_t(k)];
    if (v.empty()) continue;
    double t = 0;
    for (size_t i = 0; i < v.size(); ++i) t += v[i] * z[size_t(k) + 1 + i];
    for (size_t i = 0; i < v.size(); ++i) z[size_t(k) + 1 + i] -= 2.0 * t * v[i];
  }
  return z;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

WarmStartScan::WarmStartScan(Operator& family, hipStream_t s, const ScanOptions& o) : op_(family), s_(s), opt_(o) {
  if (!op_.parameterized()) throw invalid_options("a parameter scan needs a model with components");
  if (o.snapshots < 1 || o.snapshots > 64) throw invalid_options("snapshots must lie in [1, 64]");
  if (o.epochs < 1) throw invalid_options("epochs must be positive");
  if (!(o.learning_rate > 0.0f)) throw invalid_options("learning_rate must be positive");
  if (!(o.final_learning_rate > 0.0f && o.final_learning_rate <= o.learning_rate))
    throw invalid_options("final_learning_rate must be positive and at most learning_rate");
  P_ = int(op_.components().size());
  if (P_ > 8) throw invalid_options("the warm-start network takes at most 8 parameters");
  (void)mlp_parameter_count(P_, o.hidden1, o.hidden2, o.snapshots);
}

WarmStartScan::~WarmStartScan() = default;

// the continuation coefficients of n points, on up to 8 host threads
std::vector<std::vector<double>> WarmStartScan::continuations_of(const double* points, int n) const {
  std::vector<std::vector<double>> ys(static_cast<size_t>(std::max(n, 0)));
  const int nt = std::min(8, std::max(n, 1));
  std::vector<std::future<void>> fs;
  for (int t = 0; t < nt; ++t)
    fs.push_back(std::async(std::launch::async, [&, t] {
      for (int i = t; i < n; i += nt) ys[size_t(i)] = continuation(points + size_t(i) * P_);
    }));
  for (auto& f : fs) f.get();
  return ys;
}

std::vector<double> WarmStartScan::gaps_of(const double* points, int n) const {
  std::vector<double> g(static_cast<size_t>(std::max(n, 0)));
  const int nt = std::min(8, std::max(n, 1));
  std::vector<std::future<void>> fs;
  for (int t = 0; t < nt; ++t)
    fs.push_back(std::async(std::launch::async, [&, t] {
      for (int i = t; i < n; i += nt) (void)continuation(points + size_t(i) * P_, &g[size_t(i)]);
    }));
  for (auto& f : fs) f.get();
  return g;
}

std::vector<cplx> WarmStartScan::coefficients_of(const double* point) const {
  std::vector<cplx> c(static_cast<size_t>(P_));
  for (int p = 0; p < P_; ++p) {
    if (!std::isfinite(point[p])) throw invalid_options("scan points must be finite");
    c[size_t(p)] = cplx(point[p], 0.0);
  }
  return c;
}

// gauge of a real ground state: sign of Σ ψ_i, or of the largest entry
double WarmStartScan::gauge_sign(const double* d_psi) {
  const size_t n = op_.dim();
  const int nb = grid_for(n);
  DeviceArray<double> part(nb), amax(nb), aval(nb);
  hipLaunchKernelGGL(k_sum_blocks, dim3(nb), dim3(kT), 0, s_, d_psi, n, part.data());
  hipLaunchKernelGGL(k_absmax_blocks, dim3(nb), dim3(kT), 0, s_, d_psi, n, amax.data(), aval.data());
  std::vector<double> hp(nb), hm(nb), hv(nb);
  QMBED_HIP(hipMemcpyAsync(hp.data(), part.data(), nb * 8, hipMemcpyDeviceToHost, s_));
  QMBED_HIP(hipMemcpyAsync(hm.data(), amax.data(), nb * 8, hipMemcpyDeviceToHost, s_));
  QMBED_HIP(hipMemcpyAsync(hv.data(), aval.data(), nb * 8, hipMemcpyDeviceToHost, s_));
  QMBED_HIP(hipStreamSynchronize(s_));
  double sum = 0;
  for (double v : hp) sum += v;
  if (std::fabs(sum) > 1.0e-6) return sum < 0 ? -1.0 : 1.0;
  double bm = -1.0, bv = 1.0;
  for (int b = 0; b < nb; ++b)
    if (hm[size_t(b)] > bm) {
      bm = hm[size_t(b)];
      bv = hv[size_t(b)];
    }
  return bv < 0 ? -1.0 : 1.0;
}

// Snapshot space of the first n training states (device, N × n): the Gram
// matrix G = Ψᵀ Ψ, pivoted Cholesky of G selecting up to `snapshots` states,
// the orthonormal basis S = Ψ_sel R⁻¹ and the family's projection on it.
// Lm (n × ns, column-major) receives the coefficients of every state over S.
void WarmStartScan::build_space(const double* psi, int n, std::vector<double>& Lm, std::vector<int>& sel) {
  const size_t N = op_.dim();
  DeviceArray<double> dG(size_t(n) * size_t(n));
  if (n <= 64) {
    // few states: the tall-skinny kernel (a library GEMM of this shape may
    // run without splitting the long dimension)
    DeviceArray<double> part;
    lk::gemm_atb(psi, N, n, psi, n, dG.data(), part, s_);
    QMBED_HIP(hipStreamSynchronize(s_));
  } else {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw internal_state("rocblas_create_handle failed");
    rocblas_set_stream(h, s_);
    const double one = 1.0, zero = 0.0;
    rocblas_status st = rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, n, n, rocblas_int(N),
                                      &one, psi, rocblas_int(N), psi, rocblas_int(N), &zero, dG.data(), n);
    QMBED_HIP(hipStreamSynchronize(s_));
    rocblas_destroy_handle(h);
    if (st != rocblas_status_success) throw internal_state("rocblas_dgemm failed");
  }
  const int want = std::min(opt_.snapshots, n);
  if (opt_.space == ScanOptions::Space::Principal) {
    principal_space(psi, n, dG, want, Lm);
    sel.clear();
  } else {
    greedy_space(psi, n, dG, want, Lm, sel);
  }
  project_family();
}

// Principal components: G = U Λ Uᵀ (rocSOLVER, ascending); the leading
// `want` pairs with λ > 1e-10 λ_max give S = Ψ U_k Λ_k^{-1/2} (orthonormal)
// and the coefficients of ψ_i over S, Lm[i + n·k] = U[i][k] √λ_k.
void WarmStartScan::principal_space(const double* psi, int n, DeviceArray<double>& dG, int want,
                                    std::vector<double>& Lm) {
  const size_t N = op_.dim();
  std::vector<double> U(size_t(n) * size_t(n)), W(static_cast<size_t>(n));
  {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw internal_state("rocblas_create_handle failed");
    rocblas_set_stream(h, s_);
    DeviceArray<double> dW(static_cast<size_t>(n)), dE(static_cast<size_t>(n));
    DeviceArray<int> dInfo(1);
    rocblas_status st = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, dG.data(), n, dW.data(),
                                         dE.data(), dInfo.data());
    int info = 0;
    QMBED_HIP(hipMemcpyAsync(&info, dInfo.data(), sizeof(int), hipMemcpyDeviceToHost, s_));
    QMBED_HIP(hipMemcpyAsync(U.data(), dG.data(), U.size() * 8, hipMemcpyDeviceToHost, s_));
    QMBED_HIP(hipMemcpyAsync(W.data(), dW.data(), W.size() * 8, hipMemcpyDeviceToHost, s_));
    QMBED_HIP(hipStreamSynchronize(s_));
    rocblas_destroy_handle(h);
    if (st != rocblas_status_success || info != 0) throw internal_state("rocsolver_dsyevd failed on the Gram matrix");
  }
  std::vector<int> cols;
  for (int k = n - 1; k >= 0 && int(cols.size()) < want; --k)
    if (W[size_t(k)] > 1.0e-10 * W[size_t(n) - 1]) cols.push_back(k);
  ns_ = int(cols.size());
  if (ns_ == 0) throw internal_state("the training ground states span no snapshot");
  std::vector<double> C(size_t(n) * size_t(ns_));
  Lm.assign(size_t(n) * size_t(ns_), 0.0);
  for (int k = 0; k < ns_; ++k) {
    const double lam = W[size_t(cols[size_t(k)])];
    const double* u = U.data() + size_t(n) * size_t(cols[size_t(k)]);
    for (int i = 0; i < n; ++i) {
      C[size_t(i) + size_t(n) * size_t(k)] = u[i] / std::sqrt(lam);
      Lm[size_t(i) + size_t(n) * size_t(k)] = u[i] * std::sqrt(lam);
    }
  }
  DeviceArray<double> dC(C.size());
  QMBED_HIP(hipMemcpyAsync(dC.data(), C.data(), C.size() * 8, hipMemcpyHostToDevice, s_));
  S_.resize(N * size_t(ns_));
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) throw internal_state("rocblas_create_handle failed");
  rocblas_set_stream(h, s_);
  const double one = 1.0, zero = 0.0;
  rocblas_status st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, rocblas_int(N), ns_, n, &one,
                                    psi, rocblas_int(N), dC.data(), n, &zero, S_.data(), rocblas_int(N));
  QMBED_HIP(hipStreamSynchronize(s_));
  rocblas_destroy_handle(h);
  if (st != rocblas_status_success) throw internal_state("rocblas_dgemm failed");
}

// Pivoted Cholesky of the Gram matrix: G ≈ L Lᵀ over the selected pivots;
// S = Ψ_sel R⁻¹ with R = L[sel]ᵀ, and Lm the coefficients of ψ_i over S.
void WarmStartScan::greedy_space(const double* psi, int n, DeviceArray<double>& dG, int want,
                                 std::vector<double>& Lm, std::vector<int>& sel) {
  const size_t N = op_.dim();
  std::vector<double> G(size_t(n) * size_t(n));
  QMBED_HIP(hipMemcpy(G.data(), dG.data(), G.size() * 8, hipMemcpyDeviceToHost));
  // (column-major G)
  std::vector<double> d(static_cast<size_t>(n));
  Lm.assign(size_t(n) * size_t(want), 0.0);  // Lm[i + n*k]
  double dmax0 = 0;
  for (int i = 0; i < n; ++i) {
    d[size_t(i)] = G[size_t(i) + size_t(n) * size_t(i)];
    dmax0 = std::max(dmax0, d[size_t(i)]);
  }
  sel.clear();
  std::vector<char> used(static_cast<size_t>(n), 0);
  for (int k = 0; k < want; ++k) {
    int j = -1;
    double best = -1.0;
    for (int i = 0; i < n; ++i)
      if (!used[size_t(i)] && d[size_t(i)] > best) {
        best = d[size_t(i)];
        j = i;
      }
    if (j < 0 || best <= 1.0e-10 * dmax0) break;
    used[size_t(j)] = 1;
    sel.push_back(j);
    const double piv = std::sqrt(best);
    for (int i = 0; i < n; ++i) {
      if (used[size_t(i)] && i != j) continue;
      double v = G[size_t(i) + size_t(n) * size_t(j)];
      for (int l = 0; l < k; ++l) v -= Lm[size_t(i) + size_t(n) * size_t(l)] * Lm[size_t(j) + size_t(n) * size_t(l)];
      Lm[size_t(i) + size_t(n) * size_t(k)] = (i == j) ? piv : v / piv;
    }
    for (int i = 0; i < n; ++i) {
      if (used[size_t(i)]) continue;
      const double l = Lm[size_t(i) + size_t(n) * size_t(k)];
      d[size_t(i)] -= l * l;
    }
  }
  ns_ = int(sel.size());
  if (ns_ == 0) throw internal_state("the training ground states span no snapshot");
  // R (ns × ns, upper): R[a][b] = L[sel[b]][a]; S = Ψ_sel R⁻¹
  std::vector<double> R(size_t(ns_) * size_t(ns_), 0.0), Rinv(size_t(ns_) * size_t(ns_), 0.0);  // column-major
  for (int b = 0; b < ns_; ++b)
    for (int a = 0; a <= b; ++a) R[size_t(a) + size_t(ns_) * size_t(b)] = Lm[size_t(sel[size_t(b)]) + size_t(n) * size_t(a)];
  for (int col = 0; col < ns_; ++col) {  // solve R x = e_col (back substitution)
    for (int row = ns_ - 1; row >= 0; --row) {
      double v = (row == col) ? 1.0 : 0.0;
      for (int k = row + 1; k < ns_; ++k) v -= R[size_t(row) + size_t(ns_) * size_t(k)] * Rinv[size_t(k) + size_t(ns_) * size_t(col)];
      Rinv[size_t(row) + size_t(ns_) * size_t(col)] = v / R[size_t(row) + size_t(ns_) * size_t(row)];
    }
  }
  {
    DeviceArray<double> Psel(N * size_t(ns_)), dR(Rinv.size());
    for (int b = 0; b < ns_; ++b)
      QMBED_HIP(hipMemcpyAsync(Psel.data() + N * size_t(b), psi + N * size_t(sel[size_t(b)]), N * 8,
                               hipMemcpyDeviceToDevice, s_));
    QMBED_HIP(hipMemcpyAsync(dR.data(), Rinv.data(), Rinv.size() * 8, hipMemcpyHostToDevice, s_));
    S_.resize(N * size_t(ns_));
    lk::gemm_vc(false, Psel.data(), N, ns_, dR.data(), ns_, S_.data(), s_);
    QMBED_HIP(hipStreamSynchronize(s_));
  }
}

// Projected parts Sᵀ H_part S: H(0) S, then H(e_p) S − H(0) S per component,
// and the gauge Sᵀ1.
void WarmStartScan::project_family() {
  const size_t N = op_.dim();
  {
    DeviceArray<double> HS(N * size_t(ns_)), dP(size_t(ns_) * ns_), atb;
    proj_.assign(size_t(P_) + 1, std::vector<double>(size_t(ns_) * ns_, 0.0));
    for (int part = 0; part <= P_; ++part) {
      std::vector<cplx> c(size_t(P_), cplx(0, 0));
      if (part > 0) c[size_t(part - 1)] = cplx(1, 0);
      op_.evaluate_coefficients(c, s_);
      for (int b = 0; b < ns_; ++b) op_.apply(S_.data() + N * size_t(b), HS.data() + N * size_t(b), s_);
      lk::gemm_atb(S_.data(), N, ns_, HS.data(), ns_, dP.data(), atb, s_);
      QMBED_HIP(hipMemcpyAsync(proj_[size_t(part)].data(), dP.data(), size_t(ns_) * ns_ * 8, hipMemcpyDeviceToHost,
                               s_));
      QMBED_HIP(hipStreamSynchronize(s_));
      auto& m = proj_[size_t(part)];
      for (int i = 0; i < ns_; ++i)  // symmetrize the rounding
        for (int j = i + 1; j < ns_; ++j) {
          const double a = 0.5 * (m[size_t(j) * ns_ + i] + m[size_t(i) * ns_ + j]);
          m[size_t(j) * ns_ + i] = m[size_t(i) * ns_ + j] = a;
        }
      if (part > 0)
        for (size_t e = 0; e < m.size(); ++e) m[e] -= proj_[0][e];
    }
    // Sᵀ 1 for the gauge of continuation targets (the training states' own
    // gauge: the sign of Σ_i ψ_i)
    {
      DeviceArray<double> ones(N), dg(static_cast<size_t>(ns_));
      hipLaunchKernelGGL(k_fill, dim3(grid_for(N)), dim3(kT), 0, s_, ones.data(), N, 1.0);
      lk::gemm_atb(S_.data(), N, ns_, ones.data(), 1, dg.data(), atb, s_);
      gauge_.assign(size_t(ns_), 0.0);
      QMBED_HIP(hipMemcpyAsync(gauge_.data(), dg.data(), size_t(ns_) * 8, hipMemcpyDeviceToHost, s_));
      QMBED_HIP(hipStreamSynchronize(s_));
    }
  }
}

WarmStartScan::TrainReport WarmStartScan::train(const double* points, int n, const EigshOptions& o,
                                                const EigshControls& ctl) {
  if (n < 1) throw invalid_options("training needs at least one point");
  if (!op_.is_real() || !op_.basis().real_characters())
    throw unsupported_backend("the learned warm start is built for real-valued families");
  TrainReport rep;
  const size_t N = op_.dim();
  auto t0 = std::chrono::steady_clock::now();
  // 1. ground states of the training points: the first `snapshots` from the
  // plain start; once they are solved (and span a space), the rest start
  // from the lowest eigenvector of the family projected on that provisional
  // snapshot space, so the training solves reuse their own results
  DeviceArray<double> psi(N * size_t(n));
  const int n_plain = opt_.provisional ? std::min(n, opt_.snapshots) : n;
  bool provisional = false;
  DeviceArray<double> v0;
  rep.energies.assign(size_t(n), 0.0);
  rep.iterations.assign(size_t(n), 0);
  EigshOptions oo = o;
  oo.eigenpairs = 1;
  // one training solve on its own
  auto solve_one = [&](int i) {
    const double* point = points + size_t(i) * P_;
    const double* start = nullptr;
    if (provisional) {
      const auto y = continuation(point);
      DeviceArray<double> dy(static_cast<size_t>(ns_));
      QMBED_HIP(hipMemcpyAsync(dy.data(), y.data(), y.size() * 8, hipMemcpyHostToDevice, s_));
      lk::gemm_vc(false, S_.data(), N, ns_, dy.data(), 1, v0.data(), s_);
      QMBED_HIP(hipStreamSynchronize(s_));  // y and dy die here
      start = v0.data();
    }
    op_.evaluate_coefficients(coefficients_of(point), s_);
    if (!op_.hermitian()) throw non_hermitian();
    DeviceArray<double> vecs;
    bool cv = false;
    Eigensystem es = eigsh_on_device(op_, oo, start, false, ctl, vecs, cv, s_);
    if (cv) throw unsupported_backend("the learned warm start is built for real-valued families");
    const double sign = gauge_sign(vecs.data());
    hipLaunchKernelGGL(k_scale_copy, dim3(grid_for(N)), dim3(kT), 0, s_, vecs.data(), N, sign,
                       psi.data() + N * size_t(i));
    QMBED_HIP(hipStreamSynchronize(s_));  // vecs dies here
    rep.energies[size_t(i)] = es.eigenvalues[0];
    rep.iterations[size_t(i)] = es.iterations;
  };
  // up to `lockstep` training solves at a time in lockstep windows (the
  // same windows, products and starts as one by one, so the same states);
  // each accepted Ritz vector goes straight into its column of psi
  auto solve_batch = [&](int i0, int nb, std::vector<char>& done) {
    const int count = op_.scan_batch_amp_count();
    std::vector<std::future<std::vector<double>>> fy;
    if (provisional)
      for (int b = 0; b < nb; ++b) {
        const double* p = points + size_t(i0 + b) * P_;
        fy.push_back(std::async(std::launch::async, [this, p] { return continuation(p); }));
      }
    std::vector<std::vector<cplx>> coefs;
    for (int b = 0; b < nb; ++b) coefs.push_back(coefficients_of(points + size_t(i0 + b) * P_));
    if (lk_amps_.size() < size_t(nb) * size_t(count)) lk_amps_.resize(size_t(nb) * size_t(count));
    if (lk_diags_.size() < size_t(nb) * 2 * N) lk_diags_.resize(size_t(nb) * 2 * N);
    op_.scan_batch_tables(coefs, lk_amps_.data(), lk_diags_.data(), s_);
    std::vector<std::vector<double>> ys(static_cast<size_t>(nb));
    for (size_t b = 0; b < fy.size(); ++b) ys[b] = fy[b].get();
    std::vector<int> lanes;
    for (int b = 0; b < nb; ++b) {
      double cn = 0;
      for (double v : ys[size_t(b)]) cn += v * v;
      if (!provisional || (std::isfinite(cn) && cn > 1.0e-6)) lanes.push_back(b);
    }
    if (lanes.size() < 2) return;
    const int nl = int(lanes.size());
    if (lk_starts_.size() < size_t(nl) * N) lk_starts_.resize(size_t(nl) * N);
    if (!provisional) {
      for (int l = 0; l < nl; ++l) deterministic_start_device(lk_starts_.data() + size_t(l) * N, N, oo.seed, s_, false);
    } else {
      const size_t ny = size_t(ns_) * size_t(nl);
      cpin_.ensure(ny);
      for (int l = 0; l < nl; ++l) {
        const auto& y = ys[size_t(lanes[size_t(l)])];
        std::copy(y.begin(), y.end(), cpin_.p + size_t(l) * ns_);
      }
      if (lk_y_.size() < ny) lk_y_.resize(ny);
      QMBED_HIP(hipMemcpyAsync(lk_y_.data(), cpin_.p, ny * 8, hipMemcpyHostToDevice, s_));
      lk::gemm_vc(false, S_.data(), N, ns_, lk_y_.data(), nl, lk_starts_.data(), s_);
    }
    std::vector<LockstepPoint> pts;
    for (int l = 0; l < nl; ++l) {
      const int b = lanes[size_t(l)];
      pts.push_back({lk_starts_.data() + size_t(l) * N, lk_amps_.data() + size_t(b) * count,
                     lk_diags_.data() + size_t(b) * 2 * N, psi.data() + N * size_t(i0 + b)});
    }
    EigshControls c1 = ctl;
    c1.keep_vectors = false;  // the vectors go to psi
    const auto res = eigsh_lockstep(op_, oo, c1, pts, lk_lanes64_, lk_lanes32_, s_);
    QMBED_HIP(hipStreamSynchronize(s_));  // cpin_ is free again
    for (int l = 0; l < nl; ++l) {
      if (!res[size_t(l)].ok) continue;
      const int i = i0 + lanes[size_t(l)];
      double* col = psi.data() + N * size_t(i);
      if (gauge_sign(col) < 0) hipLaunchKernelGGL(k_scale, dim3(grid_for(N)), dim3(kT), 0, s_, col, N, -1.0);
      rep.energies[size_t(i)] = res[size_t(l)].es.eigenvalues[0];
      rep.iterations[size_t(i)] = res[size_t(l)].es.iterations;
      done[size_t(lanes[size_t(l)])] = 1;
    }
    QMBED_HIP(hipStreamSynchronize(s_));
  };
  // (small sectors take the dense route, one by one)
  // lockstep windows run the unrestarted route (explicit windows below 96
  // thick-restart, one point at a time)
  const bool batched = ctl.lockstep > 1 && N > 256 && (!o.krylov_dimension || *o.krylov_dimension >= 96) &&
                       op_.scan_batch_supported(s_);
  const int W = std::max(1, std::min(ctl.lockstep, 16));
  for (int i0 = 0; i0 < n;) {
    if (i0 == n_plain && n_plain < n) {
      std::vector<double> Lp;
      std::vector<int> sp;
      build_space(psi.data(), n_plain, Lp, sp);
      provisional = ns_ > 0;
      if (provisional) v0.resize(N);
    }
    // a batch never crosses the switch to provisional starts
    const int lim = i0 < n_plain ? n_plain : n;
    const int nb = batched ? std::min(W, lim - i0) : 1;
    std::vector<char> done(static_cast<size_t>(nb), 0);
    if (nb >= 2) solve_batch(i0, nb, done);
    for (int b = 0; b < nb; ++b)
      if (!done[size_t(b)]) solve_one(i0 + b);
    if (provisional) rep.provisional_starts += nb;
    i0 += nb;
  }
  rep.solve_seconds = seconds_since(t0);
  // 2. the snapshot space of all training states
  auto t1 = std::chrono::steady_clock::now();
  std::vector<double> Lm;
  std::vector<int> sel;
  build_space(psi.data(), n, Lm, sel);
  rep.gram_seconds = seconds_since(t1);
  // 4. the network: inputs normalized to [-1, 1] per parameter
  auto t2 = std::chrono::steady_clock::now();
  center_.assign(size_t(P_), 0.0);
  scale_.assign(size_t(P_), 1.0);
  for (int p = 0; p < P_; ++p) {
    double lo = points[size_t(p)], hi = points[size_t(p)];
    for (int i = 0; i < n; ++i) {
      lo = std::min(lo, points[size_t(i) * P_ + p]);
      hi = std::max(hi, points[size_t(i) * P_ + p]);
    }
    center_[size_t(p)] = 0.5 * (lo + hi);
    scale_[size_t(p)] = hi > lo ? 0.5 * (hi - lo) : 1.0;
  }
  // samples: the training points, then (continuation targets) seeded
  // uniform points of their bounding box
  std::vector<double> pts(points, points + size_t(n) * P_);
  if (opt_.targets == ScanOptions::Targets::Continuation) {
    Lcg rng(opt_.seed + 1);
    for (int q = 0; q < opt_.samples; ++q)
      for (int p = 0; p < P_; ++p) pts.push_back(center_[size_t(p)] + scale_[size_t(p)] * rng.next());
  }
  const int nt = int(pts.size() / size_t(P_));
  std::vector<float> X(size_t(nt) * P_), T(size_t(nt) * ns_);
  std::vector<double> Td(size_t(nt) * ns_);
  for (int i = 0; i < nt; ++i) {
    for (int p = 0; p < P_; ++p)
      X[size_t(i) * P_ + p] = float((pts[size_t(i) * P_ + p] - center_[size_t(p)]) / scale_[size_t(p)]);
    double* t = Td.data() + size_t(i) * ns_;
    if (opt_.targets == ScanOptions::Targets::Continuation) {
      const auto y = continuation(pts.data() + size_t(i) * P_);
      for (int k = 0; k < ns_; ++k) t[k] = y[size_t(k)];
    } else {
      for (int k = 0; k < ns_; ++k) t[k] = Lm[size_t(i) + size_t(n) * size_t(k)];
      if (opt_.targets == ScanOptions::Targets::Correction) {
        // the normalized projection of ψ_i (gauge as the continuation's:
        // both have a positive overlap with the all-ones vector) minus the
        // continuation coefficients at the same point
        double nn = 0;
        for (int k = 0; k < ns_; ++k) nn += t[k] * t[k];
        const double inv = nn > 0 ? 1.0 / std::sqrt(nn) : 0.0;
        const auto y = continuation(pts.data() + size_t(i) * P_);
        for (int k = 0; k < ns_; ++k) t[k] = t[k] * inv - y[size_t(k)];
      }
    }
  }
  // correction targets are small: the network learns them at unit
  // root-mean-square, and its outputs are scaled back
  target_scale_ = 1.0;
  if (opt_.targets == ScanOptions::Targets::Correction) {
    double ss = 0;
    for (double v : Td) ss += v * v;
    const double rms = Td.empty() ? 0.0 : std::sqrt(ss / double(Td.size()));
    if (rms > 0 && std::isfinite(rms)) target_scale_ = rms;
  }
  for (size_t e = 0; e < Td.size(); ++e) T[e] = float(Td[e] / target_scale_);
  const int np = mlp_parameter_count(P_, opt_.hidden1, opt_.hidden2, ns_);
  std::vector<float> init(static_cast<size_t>(np), 0.0f);
  {
    // Glorot-uniform weights, zero biases
    Lcg rng(opt_.seed);
    const int fan[3][2] = {{P_, opt_.hidden1}, {opt_.hidden1, opt_.hidden2}, {opt_.hidden2, ns_}};
    size_t off = 0;
    for (int layer = 0; layer < 3; ++layer) {
      const int fin = fan[layer][0], fout = fan[layer][1];
      const double a = std::sqrt(6.0 / double(fin + fout));
      for (int i = 0; i < fin * fout; ++i) init[off + size_t(i)] = float(a * rng.next());
      off += size_t(fin) * size_t(fout) + size_t(fout);
    }
  }
  params_.resize(static_cast<size_t>(np));
  DeviceArray<float> dX(X.size()), dT(T.size()), dLoss(1), dHist(size_t(opt_.epochs));
  QMBED_HIP(hipMemcpyAsync(params_.data(), init.data(), init.size() * 4, hipMemcpyHostToDevice, s_));
  QMBED_HIP(hipMemcpyAsync(dX.data(), X.data(), X.size() * 4, hipMemcpyHostToDevice, s_));
  QMBED_HIP(hipMemcpyAsync(dT.data(), T.data(), T.size() * 4, hipMemcpyHostToDevice, s_));
  const float decay =
      opt_.epochs > 1 ? float(std::pow(double(opt_.final_learning_rate) / double(opt_.learning_rate),
                                       1.0 / double(opt_.epochs - 1)))
                      : 1.0f;
  mlp_train_device(P_, opt_.hidden1, opt_.hidden2, ns_, params_.data(), dX.data(), dT.data(), nt, opt_.epochs,
                   opt_.learning_rate, decay, nullptr, dLoss.data(), dHist.data(), s_);
  std::vector<float> hist(size_t(opt_.epochs));
  QMBED_HIP(hipMemcpyAsync(hist.data(), dHist.data(), hist.size() * 4, hipMemcpyDeviceToHost, s_));
  QMBED_HIP(hipStreamSynchronize(s_));
  rep.train_seconds = seconds_since(t2);
  rep.loss_first = hist.front();
  rep.loss_last = hist.back();
  // relative fit of the training targets: Σ|T|² for reference
  double tn = 0;
  for (float v : T) tn += double(v) * double(v);
  rep.target_norm2 = tn / double(nt);
  rep.snapshot_indices = sel;
  rep.snapshots = ns_;
  trained_ = true;
  return rep;
}

std::vector<double> WarmStartScan::continuation(const double* point, double* gap) const {
  std::vector<double> pm = proj_[0];
  for (int p = 0; p < P_; ++p)
    for (size_t e = 0; e < pm.size(); ++e) pm[e] += point[p] * proj_[size_t(p) + 1][e];
  auto y = lowest_eigenvector(pm, ns_, gap);
  double g = 0;
  for (int k = 0; k < ns_; ++k) g += y[size_t(k)] * gauge_[size_t(k)];
  if (g < 0)
    for (double& v : y) v = -v;
  return y;
}

std::vector<double> WarmStartScan::predict(const double* points, int n) {
  if (!trained_) throw invalid_options("the warm-start network has not been trained");
  std::vector<float> X(size_t(n) * P_);
  for (int i = 0; i < n; ++i)
    for (int p = 0; p < P_; ++p) {
      if (!std::isfinite(points[size_t(i) * P_ + p])) throw invalid_options("scan points must be finite");
      X[size_t(i) * P_ + p] = float((points[size_t(i) * P_ + p] - center_[size_t(p)]) / scale_[size_t(p)]);
    }
  DeviceArray<float> dX(X.size()), dY(size_t(n) * ns_);
  QMBED_HIP(hipMemcpyAsync(dX.data(), X.data(), X.size() * 4, hipMemcpyHostToDevice, s_));
  mlp_forward_device(P_, opt_.hidden1, opt_.hidden2, ns_, params_.data(), dX.data(), n, dY.data(), s_);
  std::vector<float> Y(size_t(n) * ns_);
  QMBED_HIP(hipMemcpyAsync(Y.data(), dY.data(), Y.size() * 4, hipMemcpyDeviceToHost, s_));
  QMBED_HIP(hipStreamSynchronize(s_));
  std::vector<double> out(Y.size());
  for (size_t e = 0; e < Y.size(); ++e) out[e] = double(Y[e]) * target_scale_;
  return out;
}

std::vector<std::vector<double>> WarmStartScan::network_coefficients(const double* points, int n) {
  const auto y = predict(points, n);
  std::vector<std::vector<double>> out(static_cast<size_t>(std::max(n, 0)));
  const bool corr = opt_.targets == ScanOptions::Targets::Correction;
  std::vector<std::vector<double>> c;
  if (corr) c = continuations_of(points, n);
  for (int b = 0; b < n; ++b) {
    out[size_t(b)].assign(y.begin() + size_t(b) * ns_, y.begin() + size_t(b + 1) * ns_);
    if (corr)
      for (int k = 0; k < ns_; ++k) out[size_t(b)][size_t(k)] += c[size_t(b)][size_t(k)];
  }
  return out;
}

std::vector<WarmStartScan::PointResult> WarmStartScan::solve_many(const double* points, int n, const EigshOptions& o,
                                                                 const EigshControls& ctl, Start start) {
  std::vector<PointResult> out;
  out.reserve(size_t(std::max(n, 0)));
  if (n < 2 || (start != Start::Plain && !trained_)) {
    for (int i = 0; i < n; ++i) out.push_back(solve(points + size_t(i) * P_, o, ctl, start));
    return out;
  }
  if (ctl.lockstep > 1 && op_.dim() > 256 && (!o.krylov_dimension || *o.krylov_dimension >= 96) &&
      op_.scan_batch_supported(s_))
    return solve_lockstep(points, n, o, ctl, start);
  if (start != Start::SnapshotRR) {
    for (int i = 0; i < n; ++i) out.push_back(solve(points + size_t(i) * P_, o, ctl, start));
    return out;
  }
  auto ahead = [this](const double* p) { return std::async(std::launch::async, [this, p] { return continuation(p); }); };
  auto next = ahead(points);
  for (int i = 0; i < n; ++i) {
    const std::vector<double> y = next.get();
    if (i + 1 < n) next = ahead(points + size_t(i + 1) * P_);
    out.push_back(solve(points + size_t(i) * P_, o, ctl, start, &y));
  }
  return out;
}

// The points in lockstep (eigsh_lockstep_stream): up to ctl.lockstep
// windows step together with shared H·v launches. Free slots take the next
// points once ctl.lockstep_refill of them are free (the default, 0: when
// all are, i.e. batches that start and end together); the starts S·y of one
// refill are formed by one product, which reads S once.
// The points' amplitude tables and start coefficients are made up front
// (one upload each); a slot's diagonal is written when the slot takes its
// point. A point whose start vanishes or whose window fails is solved on its
// own.
std::vector<WarmStartScan::PointResult> WarmStartScan::solve_lockstep(const double* points, int n, const EigshOptions& o,
                                                                      const EigshControls& ctl, Start start) {
  std::vector<PointResult> out(static_cast<size_t>(n));
  const size_t N = op_.dim();
  const int count = op_.scan_batch_amp_count();
  const int W = std::min(ctl.lockstep, 16);
  EigshOptions oo = o;
  oo.eigenpairs = 1;
  EigshControls ctl1 = ctl;
  ctl1.keep_vectors = false;  // energies only
  auto t0 = std::chrono::steady_clock::now();
  auto grow = [](DeviceArray<double>& d, size_t want) {
    if (d.size() < want) d.resize(want);
  };
  // the amplitude tables of all points, in one upload
  std::vector<std::vector<cplx>> coefs(static_cast<size_t>(n));
  amp_pin_.ensure(size_t(n) * size_t(count));
  for (int i = 0; i < n; ++i) {
    coefs[size_t(i)] = coefficients_of(points + size_t(i) * P_);
    const auto a = op_.scan_amp_table(coefs[size_t(i)]);
    std::copy(a.begin(), a.end(), amp_pin_.p + size_t(i) * count);
  }
  grow(lk_amps_, size_t(n) * size_t(count));
  QMBED_HIP(hipMemcpyAsync(lk_amps_.data(), amp_pin_.p, size_t(n) * size_t(count) * 8, hipMemcpyHostToDevice, s_));
  // the start coefficients of all points: the continuation (host threads)
  // or the network's (one forward pass); the plain start needs none
  std::vector<std::vector<double>> ys;
  if (start == Start::SnapshotRR) ys = continuations_of(points, n);
  else if (start == Start::Network) ys = network_coefficients(points, n);
  std::vector<int> queue;  // points with a usable start, in order
  for (int i = 0; i < n; ++i) {
    double cn = 0;
    if (start != Start::Plain)
      for (double v : ys[size_t(i)]) cn += v * v;
    if (start == Start::Plain || (std::isfinite(cn) && cn > 1.0e-6)) queue.push_back(i);
  }
  // warm starts: the hardest points first (ascending projected gap), so a
  // batch's windows need similar step counts; ties keep the given order
  if (start != Start::Plain && ctl.lockstep_order && queue.size() > 1) {
    const auto gaps = gaps_of(points, n);
    std::stable_sort(queue.begin(), queue.end(), [&](int a, int b) { return gaps[size_t(a)] < gaps[size_t(b)]; });
  }
  // the usable points' coefficients in queue order, in one upload
  const size_t nq = queue.size();
  if (start != Start::Plain && nq > 0) {
    cpin_.ensure(nq * ns_);
    for (size_t q = 0; q < nq; ++q) std::copy(ys[size_t(queue[q])].begin(), ys[size_t(queue[q])].end(), cpin_.p + q * ns_);
    grow(lk_y_, nq * ns_);
    QMBED_HIP(hipMemcpyAsync(lk_y_.data(), cpin_.p, nq * ns_ * 8, hipMemcpyHostToDevice, s_));
  }
  std::vector<char> solved(static_cast<size_t>(n), 0);
  int solved_n = 0;
  LockstepFeed feed;
  feed.done = [&](int i, LockstepResult&& r) {
    if (!r.ok) return;
    PointResult& pr = out[size_t(i)];
    pr.es = std::move(r.es);
    pr.warm = start != Start::Plain;
    solved[size_t(i)] = 1;
    ++solved_n;
  };
  if (std::min<size_t>(size_t(W), nq) >= 2) {
    const int slots = int(std::min<size_t>(size_t(W), nq));
    grow(lk_diags_, size_t(slots) * 2 * N);
    grow(lk_starts_, size_t(slots) * N);
    size_t cursor = 0, q0 = 0;
    feed.group = ctl.lockstep_refill <= 0 || ctl.lockstep_refill >= slots ? slots : ctl.lockstep_refill;
    // the starts of one refill (up to `slots` points), formed together: S is
    // read once per refill; a slot's diagonal goes into the slot's own area
    feed.prepare = [&](int nfree) {
      q0 = cursor;
      const int cnt = int(std::min<size_t>(size_t(nfree), nq - cursor));
      if (cnt <= 0) return;
      if (start == Start::Plain) {
        for (int b = 0; b < cnt; ++b) deterministic_start_device(lk_starts_.data() + size_t(b) * N, N, oo.seed, s_, false);
      } else {
        lk::gemm_vc(false, S_.data(), N, ns_, lk_y_.data() + q0 * ns_, cnt, lk_starts_.data(), s_);
      }
    };
    feed.next = [&](int slot, LockstepPoint& pt) {
      if (cursor >= nq) return -1;
      const size_t q = cursor++;
      const int i = queue[q];
      double* diag = lk_diags_.data() + size_t(slot) * 2 * N;
      op_.scan_point_diagonal(coefs[size_t(i)], diag, s_);
      pt = LockstepPoint{lk_starts_.data() + (q - q0) * N, lk_amps_.data() + size_t(i) * count, diag};
      return i;
    };
    eigsh_lockstep_stream(op_, oo, ctl1, slots, feed, lk_lanes64_, lk_lanes32_, s_);
  }
  QMBED_HIP(hipStreamSynchronize(s_));  // the pinned staging areas are free again
  const double dt = seconds_since(t0);
  for (int i = 0; i < n; ++i)
    if (solved[size_t(i)]) out[size_t(i)].seconds = dt / double(std::max(solved_n, 1));
  for (int i = 0; i < n; ++i)
    if (!solved[size_t(i)])
      out[size_t(i)] = solve(points + size_t(i) * P_, o, ctl, start,
                             start == Start::SnapshotRR ? &ys[size_t(i)] : nullptr);
  return out;
}

WarmStartScan::PointResult WarmStartScan::solve(const double* point, const EigshOptions& o, const EigshControls& ctl_in,
                                                Start start, const std::vector<double>* y) {
  EigshControls ctl = ctl_in;
  ctl.keep_vectors = false;  // energies only
  const bool warm = start != Start::Plain;
  PointResult r;
  auto t0 = std::chrono::steady_clock::now();
  op_.evaluate_coefficients(coefficients_of(point), s_);
  if (!op_.hermitian()) throw non_hermitian();
  EigshOptions oo = o;
  oo.eigenpairs = 1;
  DeviceArray<double> vecs;
  bool cv = false;
  const size_t N = op_.dim();
  if (warm) {
    if (!trained_) throw invalid_options("the warm-start network has not been trained");
    // coefficients y (network on the device, or the projected family on the
    // host); the start vector v₀ = S y
    if (v0_.size() != N) v0_.resize(N);
    if (dc_.size() != size_t(ns_)) dc_.resize(static_cast<size_t>(ns_));
    std::vector<double> c(static_cast<size_t>(ns_));
    const bool host_coefficients =
        start == Start::SnapshotRR || opt_.targets == ScanOptions::Targets::Correction;
    if (host_coefficients) {
      c = start == Start::SnapshotRR ? (y ? *y : continuation(point)) : network_coefficients(point, 1)[0];
      if (c.size() != size_t(ns_)) throw internal_state("continuation coefficients of the wrong length");
      // staged through pinned memory (the previous point's copy completed
      // before its solve returned)
      cpin_.ensure(c.size());
      std::copy(c.begin(), c.end(), cpin_.p);
      QMBED_HIP(hipMemcpyAsync(dc_.data(), cpin_.p, c.size() * 8, hipMemcpyHostToDevice, s_));
      lk::gemm_vc(false, S_.data(), N, ns_, dc_.data(), 1, v0_.data(), s_);
    } else {
      std::vector<float> x(static_cast<size_t>(P_));
      for (int p = 0; p < P_; ++p) x[size_t(p)] = float((point[p] - center_[size_t(p)]) / scale_[size_t(p)]);
      if (dx_.size() != size_t(P_)) dx_.resize(static_cast<size_t>(P_));
      if (dy_.size() != size_t(ns_)) dy_.resize(static_cast<size_t>(ns_));
      QMBED_HIP(hipMemcpyAsync(dx_.data(), x.data(), x.size() * 4, hipMemcpyHostToDevice, s_));
      mlp_forward_device(P_, opt_.hidden1, opt_.hidden2, ns_, params_.data(), dx_.data(), 1, dy_.data(), s_);
      hipLaunchKernelGGL(k_f2d, dim3(1), dim3(64), 0, s_, dy_.data(), ns_, dc_.data());
      lk::gemm_vc(false, S_.data(), N, ns_, dc_.data(), 1, v0_.data(), s_);
      QMBED_HIP(hipMemcpyAsync(c.data(), dc_.data(), c.size() * 8, hipMemcpyDeviceToHost, s_));
      QMBED_HIP(hipStreamSynchronize(s_));  // c (host) and x die here
    }
    double cn = 0;
    for (double v : c) cn += v * v;
    // S is orthonormal, so ‖v₀‖ = ‖y‖; a vanishing or non-finite prediction
    // falls back to the plain start
    if (std::isfinite(cn) && cn > 1.0e-6) {
      r.es = eigsh_on_device(op_, oo, v0_.data(), false, ctl, vecs, cv, s_);
      r.warm = true;
    } else {
      r.es = eigsh_on_device(op_, oo, nullptr, false, ctl, vecs, cv, s_);
      r.fallback = true;
    }
  } else {
    r.es = eigsh_on_device(op_, oo, nullptr, false, ctl, vecs, cv, s_);
  }
  r.seconds = seconds_since(t0);
  return r;
}

}  // namespace qmbed
