// Shift-invert eigsh for the Shift target on gfx950.
//
// Restates the reference's route for a matrix-free operator (QMBED
// src/solve/mod.rs:950-979 transformed_apply, :1285-1456 lanczos_eigsh,
// :2242-2296 expanding_lanczos_eigsh): a fully reorthogonalized Lanczos
// window on the transformed operator (A − σI)⁻¹, the Ritz values of largest
// magnitude selected from its tridiagonal matrix, every Ritz vector's
// Rayleigh quotient and residual taken with A itself, the pairs ordered by
// |λ − σ| and accepted when each residual is at most tol·max(|λ|, 1); a
// failed window doubles within the iteration budget, and the full window of
// at most 2048 rows falls back to the dense route.
//
// The inner solves (A − σI) x = b run MINRES (Paige and Saunders, 1975) on
// the device: (A − σI) is Hermitian and usually indefinite, MINRES minimizes
// the residual over the same Krylov space as the reference's restarted
// GMRES (src/solve/mod.rs:760-851) with a three-term recurrence and no
// stored basis. Each solve stops at the reference's relative residual
// min(0.1·tol, 1e-10) and is confirmed by a true residual (a MINRES restart
// from the current solution otherwise), within max(max_iterations, 128)
// products; a solve that does not get there fails as NonConvergence, as the
// reference's GMRES does. Vectors are complex FP64 (the reference's complex
// path; its real shift-invert path needs a sparse factorization).
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace qmbed {

Eigensystem dense_eigsh_route(const Operator& op, const EigshOptions& o, DeviceArray<double>& vec_out,
                              bool& complex_out, hipStream_t s);

namespace {

constexpr int kT = 256;
constexpr int kParts = 256;  // blocks of the reductions (fixed-order partials)

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kT];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  return red[0];
}

// partials of Re⟨a, b⟩ (b == nullptr: ‖a‖²)
__global__ void __launch_bounds__(kT) si_dot(const double2* __restrict__ a, const double2* __restrict__ b, size_t n,
                                             double* part) {
  double acc = 0;
  for (size_t i = size_t(blockIdx.x) * kT + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double2 x = a[i], y = b ? b[i] : a[i];
    acc += x.x * y.x + x.y * y.y;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// y = av − σ v − c1 r1 (r1 may alias y), partials of Re⟨v, y⟩
__global__ void __launch_bounds__(kT) si_shifted(const double2* __restrict__ av, const double2* __restrict__ v,
                                                 const double2* r1, double sigma, double c1, size_t n, double2* y,
                                                 double* part) {
  double acc = 0;
  for (size_t i = size_t(blockIdx.x) * kT + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double2 a = av[i], x = v[i];
    double2 t = make_double2(a.x - sigma * x.x, a.y - sigma * x.y);
    if (c1 != 0.0) {
      const double2 r = r1[i];
      t.x -= c1 * r.x;
      t.y -= c1 * r.y;
    }
    y[i] = t;
    acc += x.x * t.x + x.y * t.y;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// y −= c r, partials of ‖y‖²
__global__ void __launch_bounds__(kT) si_sub_norm(double2* y, const double2* __restrict__ r, double c, size_t n,
                                                  double* part) {
  double acc = 0;
  for (size_t i = size_t(blockIdx.x) * kT + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 t = y[i];
    const double2 q = r[i];
    t.x -= c * q.x;
    t.y -= c * q.y;
    y[i] = t;
    acc += t.x * t.x + t.y * t.y;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out = s·in
__global__ void __launch_bounds__(kT) si_scale(const double2* __restrict__ in, double s, size_t n, double2* out) {
  for (size_t i = size_t(blockIdx.x) * kT + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double2 t = in[i];
    out[i] = make_double2(s * t.x, s * t.y);
  }
}

// wn = (v − e·wp − d·wc)·g (wn may alias neither wp nor wc), x += φ·wn
__global__ void __launch_bounds__(kT) si_wstep(const double2* __restrict__ v, const double2* __restrict__ wp,
                                               const double2* __restrict__ wc, double e, double d, double g, double phi,
                                               size_t n, double2* wn, double2* x) {
  for (size_t i = size_t(blockIdx.x) * kT + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double2 a = v[i], p = wp[i], c = wc[i];
    const double2 t = make_double2((a.x - e * p.x - d * c.x) * g, (a.y - e * p.y - d * c.y) * g);
    wn[i] = t;
    double2 xv = x[i];
    xv.x += phi * t.x;
    xv.y += phi * t.y;
    x[i] = xv;
  }
}

// r = b − (av − σ x), partials of ‖r‖²
__global__ void __launch_bounds__(kT) si_residual(const double2* __restrict__ b, const double2* __restrict__ ax,
                                                  const double2* __restrict__ x, double sigma, size_t n, double2* r,
                                                  double* part) {
  double acc = 0;
  for (size_t i = size_t(blockIdx.x) * kT + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    const double2 q = b[i], a = ax[i], u = x[i];
    const double2 t = make_double2(q.x - (a.x - sigma * u.x), q.y - (a.y - sigma * u.y));
    r[i] = t;
    acc += t.x * t.x + t.y * t.y;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// w −= α v + β u (u may be null)
__global__ void __launch_bounds__(kT) si_three(double2* w, const double2* __restrict__ v, const double2* __restrict__ u,
                                               double alpha, double beta, size_t n) {
  for (size_t i = size_t(blockIdx.x) * kT + threadIdx.x; i < n; i += size_t(gridDim.x) * kT) {
    double2 t = w[i];
    const double2 a = v[i];
    t.x -= alpha * a.x;
    t.y -= alpha * a.y;
    if (u) {
      const double2 b = u[i];
      t.x -= beta * b.x;
      t.y -= beta * b.y;
    }
    w[i] = t;
  }
}

// a real start into complex interleaved
__global__ void __launch_bounds__(kT) si_to_complex(const double* __restrict__ in, size_t n, double2* out) {
  for (size_t i = size_t(blockIdx.x) * kT + threadIdx.x; i < n; i += size_t(gridDim.x) * kT)
    out[i] = make_double2(in[i], 0.0);
}

struct Solver {
  const Operator& op;
  double sigma;
  size_t n;
  hipStream_t s;
  unsigned grid;
  DeviceArray<double2> v, av, r1, r2, w0, w1, w2, rr;
  DeviceArray<double> part;
  std::vector<double> hp;
  long products = 0;  // H·v launched by the inner solves

  Solver(const Operator& o, double sg, hipStream_t st) : op(o), sigma(sg), n(o.dim()), s(st) {
    grid = unsigned(std::min<size_t>(kParts, (n + kT - 1) / kT));
    for (auto* a : {&v, &av, &r1, &r2, &w0, &w1, &w2, &rr}) a->resize(n);
    part.resize(kParts);
    hp.resize(kParts);
  }
  // the fixed-order sum of the partials of the last reduction launch
  double sum() {
    QMBED_HIP(hipMemcpyAsync(hp.data(), part.data(), grid * sizeof(double), hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
    double t = 0;
    for (unsigned i = 0; i < grid; ++i) t += hp[i];
    return t;
  }
  double norm(const double2* a) {
    hipLaunchKernelGGL(si_dot, dim3(grid), dim3(kT), 0, s, a, nullptr, n, part.data());
    return std::sqrt(sum());
  }
  void apply(const double2* in, double2* out) {
    op.apply(reinterpret_cast<const cplx*>(in), reinterpret_cast<cplx*>(out), s);
    ++products;
  }

  // x = (A − σI)⁻¹ b to ‖b − (A − σI)x‖ ≤ tol‖b‖; `budget` products
  void solve(const double2* b, double2* x, double tol, long budget) {
    const double bnorm = norm(b);
    QMBED_HIP(hipMemsetAsync(x, 0, n * sizeof(double2), s));
    if (bnorm <= 2.220446049250313e-16) return;
    long spent = 0;
    const double2* rhs = b;  // right-hand side of the current MINRES run
    double rnorm = bnorm;
    for (;;) {
      // MINRES on (A − σI) d = rhs, x += d
      const double beta1 = rnorm;
      double oldb = 0, beta = beta1, dbar = 0, epsln = 0, phibar = beta1, cs = -1, sn = 0;
      QMBED_HIP(hipMemsetAsync(w1.data(), 0, n * sizeof(double2), s));
      QMBED_HIP(hipMemsetAsync(w2.data(), 0, n * sizeof(double2), s));
      double2 *pr1 = r1.data(), *pr2 = r2.data();
      double2 *wp = w1.data(), *wc = w2.data(), *wf = w0.data();
      QMBED_HIP(hipMemcpyAsync(pr2, rhs, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
      bool first = true;
      while (spent < budget) {
        // v = r2 / β; y = (A − σ)v − (β/oldb) r1, written over r1
        hipLaunchKernelGGL(si_scale, dim3(grid), dim3(kT), 0, s, pr2, 1.0 / beta, n, v.data());
        apply(v.data(), av.data());
        ++spent;
        hipLaunchKernelGGL(si_shifted, dim3(grid), dim3(kT), 0, s, av.data(), v.data(), pr1, sigma,
                           first ? 0.0 : beta / oldb, n, pr1, part.data());
        const double alfa = sum();
        hipLaunchKernelGGL(si_sub_norm, dim3(grid), dim3(kT), 0, s, pr1, pr2, alfa / beta, n, part.data());
        oldb = beta;
        beta = std::sqrt(sum());
        std::swap(pr1, pr2);  // r1 <- old r2, r2 <- y
        const double oldeps = epsln;
        const double delta = cs * dbar + sn * alfa;
        const double gbar = sn * dbar - cs * alfa;
        epsln = sn * beta;
        dbar = -cs * beta;
        double gamma = std::hypot(gbar, beta);
        gamma = std::max(gamma, 2.220446049250313e-16);
        cs = gbar / gamma;
        sn = beta / gamma;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        // w_new = (v − oldeps w_{k−2} − delta w_{k−1}) / gamma, x += φ w_new
        hipLaunchKernelGGL(si_wstep, dim3(grid), dim3(kT), 0, s, v.data(), wp, wc, oldeps, delta, 1.0 / gamma, phi,
                           n, wf, x);
        double2* t = wp;
        wp = wc;
        wc = wf;
        wf = t;
        first = false;
        if (phibar <= 0.5 * tol * bnorm || beta <= 1.0e-300) break;
      }
      // the true residual decides
      apply(x, av.data());
      ++spent;
      hipLaunchKernelGGL(si_residual, dim3(grid), dim3(kT), 0, s, b, av.data(), x, sigma, n, rr.data(), part.data());
      rnorm = std::sqrt(sum());
      if (rnorm <= tol * bnorm) return;
      if (spent >= budget) throw non_convergence(spent, rnorm);
      // restart on the true residual: the next run adds (A − σI)⁻¹ r to x
      // (rr is not written during a run)
      rhs = rr.data();
    }
  }
};

struct Candidate {
  double value;
  double residual;
  int order;
};

// one window: Lanczos on (A − σI)⁻¹ of at most m steps from `start`
// (normalized complex); returns the accepted pairs or throws NonConvergence
Eigensystem si_window(const Operator& op, const EigshOptions& o, int m, const double2* start,
                      DeviceArray<double>& vec_out, hipStream_t s, long inner_budget, double inner_tol) {
  const size_t n = op.dim();
  Solver sv(op, o.shift, s);
  DeviceArray<double2> V(n * size_t(m)), w(n), au(n);
  DeviceArray<double2> h(static_cast<size_t>(m)), h2(static_cast<size_t>(m));
  rocblas_handle hb;
  if (rocblas_create_handle(&hb) != rocblas_status_success) throw internal_state("rocblas_create_handle failed");
  rocblas_set_stream(hb, s);
  rocblas_set_pointer_mode(hb, rocblas_pointer_mode_host);
  struct Guard {
    rocblas_handle h;
    ~Guard() { rocblas_destroy_handle(h); }
  } guard{hb};
  const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0}, mone{-1.0, 0.0};
  auto rb = [](const double2* p) { return reinterpret_cast<const rocblas_double_complex*>(p); };
  auto rbw = [](double2* p) { return reinterpret_cast<rocblas_double_complex*>(p); };
  auto check = [](rocblas_status st) {
    if (st != rocblas_status_success) throw internal_state("rocblas call failed in the shift-invert Lanczos window");
  };
  QMBED_HIP(hipMemcpyAsync(V.data(), start, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
  std::vector<double> alpha, beta;
  const unsigned grid = sv.grid;
  long steps = 0;
  // the accepted eigenpairs of the window's first `size` columns, or the
  // largest failing residual
  auto extract = [&](int size, bool final_check, Eigensystem& es, double& fail) -> bool {
    std::vector<double> a(alpha.begin(), alpha.begin() + size), b(beta.begin(), beta.begin() + (size - 1));
    std::vector<double> vals, vecs;
    tridiagonal_eigen(a, b, vals, &vecs);
    const auto sel = select_indices(vals, SpectrumTarget::LargestMagnitude, o.eigenpairs, 0.0);
    std::vector<Candidate> cand;
    DeviceArray<double2> Y(static_cast<size_t>(size));
    std::vector<double2> yh(static_cast<size_t>(size));
    DeviceArray<double> vecs_dev(2 * n * sel.size());
    for (size_t q = 0; q < sel.size(); ++q) {
      for (int r = 0; r < size; ++r) yh[size_t(r)] = make_double2(vecs[size_t(r) * size + size_t(sel[q])], 0.0);
      QMBED_HIP(hipMemcpyAsync(Y.data(), yh.data(), size_t(size) * sizeof(double2), hipMemcpyHostToDevice, s));
      double2* uq = reinterpret_cast<double2*>(vecs_dev.data()) + n * q;
      check(rocblas_zgemv(hb, rocblas_operation_none, rocblas_int(n), rocblas_int(size), &one, rb(V.data()),
                          rocblas_int(n), rb(Y.data()), 1, &zero, rbw(uq), 1));
      const double nu = sv.norm(uq);
      if (!(nu > 0)) throw non_convergence(size, INFINITY);
      hipLaunchKernelGGL(si_scale, dim3(grid), dim3(kT), 0, s, uq, 1.0 / nu, n, uq);
      sv.apply(uq, au.data());
      // Rayleigh quotient and residual ‖Au − λu‖ with A itself
      hipLaunchKernelGGL(si_dot, dim3(grid), dim3(kT), 0, s, uq, au.data(), n, sv.part.data());
      const double lam = sv.sum();
      hipLaunchKernelGGL(si_three, dim3(grid), dim3(kT), 0, s, au.data(), uq, nullptr, lam, 0.0, n);
      const double res = sv.norm(au.data());
      cand.push_back({lam, res, int(q)});
    }
    // ordered by |λ − σ| (a stable sort keeps the selection order on ties)
    std::stable_sort(cand.begin(), cand.end(), [&](const Candidate& x, const Candidate& y) {
      return std::fabs(x.value - o.shift) < std::fabs(y.value - o.shift);
    });
    fail = 0;
    for (auto& c : cand)
      if (c.residual > o.tolerance * std::max(std::fabs(c.value), 1.0)) fail = std::max(fail, c.residual);
    (void)final_check;
    if (fail > 0) return false;
    es = Eigensystem();
    vec_out.resize(2 * n * cand.size());
    for (size_t q = 0; q < cand.size(); ++q) {
      es.eigenvalues.push_back(cand[q].value);
      es.residuals.push_back(cand[q].residual);
      QMBED_HIP(hipMemcpyAsync(vec_out.data() + 2 * n * q, vecs_dev.data() + 2 * n * size_t(cand[q].order),
                               2 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    QMBED_HIP(hipStreamSynchronize(s));
    es.iterations = size;
    es.converged = true;
    return true;
  };

  int size = 0;
  for (int j = 0; j < m; ++j) {
    double2* vj = V.data() + n * size_t(j);
    sv.solve(vj, w.data(), inner_tol, inner_budget);
    ++steps;
    hipLaunchKernelGGL(si_dot, dim3(grid), dim3(kT), 0, s, vj, w.data(), n, sv.part.data());
    const double a = sv.sum();
    alpha.push_back(a);
    hipLaunchKernelGGL(si_three, dim3(grid), dim3(kT), 0, s, w.data(), vj, j ? vj - n : nullptr, a,
                       j ? beta[size_t(j - 1)] : 0.0, n);
    // classical Gram-Schmidt against the window with the DGKS criterion
    const double w0 = sv.norm(w.data());
    check(rocblas_zgemv(hb, rocblas_operation_conjugate_transpose, rocblas_int(n), rocblas_int(j + 1), &one,
                        rb(V.data()), rocblas_int(n), rb(w.data()), 1, &zero, rbw(h.data()), 1));
    check(rocblas_zgemv(hb, rocblas_operation_none, rocblas_int(n), rocblas_int(j + 1), &mone, rb(V.data()),
                        rocblas_int(n), rb(h.data()), 1, &one, rbw(w.data()), 1));
    double bnorm = sv.norm(w.data());
    if (bnorm <= 0.7071067811865476 * w0) {
      check(rocblas_zgemv(hb, rocblas_operation_conjugate_transpose, rocblas_int(n), rocblas_int(j + 1), &one,
                          rb(V.data()), rocblas_int(n), rb(w.data()), 1, &zero, rbw(h2.data()), 1));
      check(rocblas_zgemv(hb, rocblas_operation_none, rocblas_int(n), rocblas_int(j + 1), &mone, rb(V.data()),
                          rocblas_int(n), rb(h2.data()), 1, &one, rbw(w.data()), 1));
      bnorm = sv.norm(w.data());
    }
    size = j + 1;
    if (j + 1 == m || bnorm <= 1.0e-14) break;
    beta.push_back(bnorm);
    hipLaunchKernelGGL(si_scale, dim3(grid), dim3(kT), 0, s, w.data(), 1.0 / bnorm, n, V.data() + n * size_t(j + 1));
    // early acceptance: every 16 steps once the window holds more vectors
    // than pairs, the FP64 residual test with A decides
    if (size > o.eigenpairs && size % 16 == 0) {
      Eigensystem es;
      double fail = 0;
      if (extract(size, false, es, fail)) return es;
    }
  }
  if (size <= o.eigenpairs) throw non_convergence(size, INFINITY);
  Eigensystem es;
  double fail = 0;
  if (!extract(size, true, es, fail)) throw non_convergence(size, fail);
  return es;
}

}  // namespace

Eigensystem shift_invert_eigsh(const Operator& op, const EigshOptions& o, const double* d_initial,
                               bool initial_complex, DeviceArray<double>& vec_out, bool& complex_vectors,
                               hipStream_t s) {
  const size_t n = op.dim();
  if (!op.hermitian()) throw non_hermitian();
  if (o.krylov_dimension && *o.krylov_dimension < 96)
    throw unsupported_backend("shift-invert targets with Krylov windows below 96 (the thick restart) are not built on "
                              "the MI355X path");
  complex_vectors = true;
  // the complex normalized start
  DeviceArray<double2> start(n);
  const unsigned grid = unsigned(std::min<size_t>(kParts, (n + kT - 1) / kT));
  if (d_initial) {
    if (initial_complex) {
      QMBED_HIP(hipMemcpyAsync(start.data(), d_initial, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    } else {
      hipLaunchKernelGGL(si_to_complex, dim3(grid), dim3(kT), 0, s, d_initial, n, start.data());
    }
  } else {
    deterministic_start_device(reinterpret_cast<double*>(start.data()), n, o.seed, s, true);
  }
  {
    DeviceArray<double> part(kParts);
    std::vector<double> hp(kParts);
    hipLaunchKernelGGL(si_dot, dim3(grid), dim3(kT), 0, s, start.data(), nullptr, n, part.data());
    QMBED_HIP(hipMemcpyAsync(hp.data(), part.data(), grid * sizeof(double), hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
    double t = 0;
    for (unsigned i = 0; i < grid; ++i) t += hp[i];
    const double nrm = std::sqrt(t);
    if (!(nrm > 0) || !std::isfinite(nrm)) throw non_convergence(0, nrm);
    hipLaunchKernelGGL(si_scale, dim3(grid), dim3(kT), 0, s, start.data(), 1.0 / nrm, n, start.data());
  }
  // transformed_apply's inner tolerance and product budget
  const double inner_tol = std::min(0.1 * o.tolerance, 1.0e-10);
  const long inner_budget = std::max<long>(o.max_iterations, 128);
  // expanding_lanczos_eigsh: window doubling within the budget, dense
  // fallback for a full window of at most 2048 rows
  const bool fixed = bool(o.krylov_dimension);
  long sub = fixed ? std::min<long>(std::min<long>(*o.krylov_dimension, o.max_iterations), long(n))
                   : std::min<long>(std::min<long>(std::max<long>(8L * o.eigenpairs + 64, 256), long(n)),
                                    o.max_iterations);
  if (sub <= o.eigenpairs) throw invalid_options("the effective Krylov dimension must exceed eigenpairs");
  long spent = 0;
  for (;;) {
    try {
      Eigensystem es = si_window(op, o, int(sub), start.data(), vec_out, s, inner_budget, inner_tol);
      es.iterations += spent;
      return es;
    } catch (Error& err) {
      if (err.kind != ErrorKind::NonConvergence || fixed) throw;
      spent += err.iterations;
      if (sub == long(n) && n <= 2048) {
        bool c = false;
        Eigensystem es = dense_eigsh_route(op, o, vec_out, c, s);
        es.iterations += spent;
        complex_vectors = c;
        return es;
      }
      const long remaining = o.max_iterations > spent ? o.max_iterations - spent : 0;
      const long next = std::min(std::min(sub * 2, long(n)), remaining);
      if (next <= sub) throw non_convergence(spent, err.residual);
      sub = next;
    }
  }
}

}  // namespace qmbed
