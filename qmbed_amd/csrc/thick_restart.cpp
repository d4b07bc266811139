// Thick-restarted Lanczos eigensolver on device vectors: QMBED
// src/solve/mod.rs:1644-1649 (thick_restart_dimension), :1651-1913
// (thick_restarted_eigsh_real), :1914-2167 (the complex variant), :2169-2196
// (the real/complex choice) and the dense fallback of :2306-2327.
//
// A cycle keeps up to half the window of Ritz vectors (with their H·v, so
// restarts cost no products), extends them by the residual directions and a
// Krylov sequence until the window is full, and takes the Rayleigh-Ritz
// pairs of the window from the overlaps its orthonormalization recorded.
// Converged pairs are locked: later vectors are kept orthogonal to them.
// The host drives the loop (a cycle is a few dozen steps); every vector
// operation runs on the device: H·v, the overlaps and subtractions of the
// Gram-Schmidt passes (classical Gram-Schmidt against the locked vectors and
// the window at once, the reference's DGKS criterion for a second pass), the
// Ritz assembly V·Y, norms, Rayleigh quotients and residuals. The projected
// matrix (window order) is solved on the host. Matvecs are FP64.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "internal.hpp"
#include "lanczos_kernels.hpp"

namespace qmbed {

Eigensystem dense_eigsh_route(const Operator& op, const EigshOptions& o, DeviceArray<double>& vec_out,
                              bool& complex_out, hipStream_t s);

namespace {

constexpr double kDgks = 0.7071067811865476;  // DGKS_REORTHOGONALIZATION_THRESHOLD
constexpr double kEps = 2.220446049250313e-16;

bool target_before(SpectrumTarget t, double a, double b) {
  switch (t) {
    case SpectrumTarget::LargestAlgebraic: return b < a;
    case SpectrumTarget::SmallestMagnitude: return std::fabs(a) < std::fabs(b);
    case SpectrumTarget::LargestMagnitude: return std::fabs(b) < std::fabs(a);
    default: return a < b;  // smallest algebraic, both ends
  }
}

class ThickRestart {
 public:
  ThickRestart(const Operator& op, const EigshOptions& o, bool c, hipStream_t s)
      : op_(op), o_(o), c_(c), n_(op.dim()), e_(c ? 2 : 1), s_(s) {}

  Eigensystem run(const std::vector<const double*>& init, DeviceArray<double>& vec_out) {
    const long k = o_.eigenpairs;
    rd_ = std::min<long>(o_.krylov_dimension ? *o_.krylov_dimension : std::max<long>(4 * k + 24, 32), long(n_));
    if (rd_ <= k) throw invalid_options("the effective Krylov dimension must exceed eigenpairs");
    const long keep_max = std::max<long>(std::min<long>(k, rd_ / 2), 1);
    const size_t col = e_ * n_;
    B_.resize(size_t(k + rd_) * col);
    A_.resize(size_t(rd_) * col);
    R_.resize(size_t(keep_max) * col);
    RA_.resize(size_t(keep_max) * col);
    Cv_.resize(size_t(keep_max) * col);
    Ca_.resize(size_t(keep_max) * col);
    w_.resize(col);
    h_.resize(size_t(k + rd_ + 1) * 2);
    sl_.resize(size_t(4 * keep_max + 4));
    U_.resize(size_t(rd_) * size_t(keep_max) * 2);
    pin_.ensure(size_t(k + rd_ + 1) * 2 + size_t(4 * keep_max + 4));

    long L = 0;  // locked pairs, in B[0, L)
    std::vector<double> locked_value, locked_residual;
    long iterations = 0, cycle = 0, nret = 0;
    double last_residual = INFINITY;
    bool first = true;
    while (iterations < o_.max_iterations && L < k) {
      long j = 0, applied = 0;  // window vectors in B[L, L + j); their H·v in A[0, applied)
      const bool reuse = !first && nret > 0;
      if (first) {
        for (const double* v : init) {
          copy(w_.data(), v);
          if (orthonormalize(L + j, L, Bc(L + j), nullptr)) ++j;
          if (j >= rd_ / 2) break;
        }
      } else if (reuse) {
        for (long r = 0; r < nret; ++r) {
          lk::nrm2sq_to_slot(c_, Rc(r), n_, sl_.data(), ws_, s_);
          const double nr = std::sqrt(read(sl_.data(), 1)[0]);
          if (std::isfinite(nr) && nr > 1e-14) {
            lk::normalize(c_, Rc(r), sl_.data(), n_, Bc(L + j), nullptr, nullptr, s_);
            lk::normalize(c_, RAc(r), sl_.data(), n_, Ac(applied), nullptr, nullptr, s_);
            ++j;
            ++applied;
          }
          if (j >= rd_ / 2) break;
        }
      }
      first = false;
      if (j == 0) {
        deterministic_start_device(w_.data(), n_, o_.seed + uint64_t(cycle) * 0x9e3779b9ULL, s_, c_);
        if (orthonormalize(L, L, Bc(L), nullptr)) j = 1;
      }
      if (j == 0) break;
      const long known = applied;
      std::vector<std::vector<cplx>> columns;
      std::vector<cplx> proj;
      if (reuse) {
        for (long cc = 0; cc < known; ++cc) {
          if (j >= rd_) break;
          copy(w_.data(), Ac(cc));
          const bool ok = orthonormalize(L + j, L, Bc(L + j), &proj);
          columns.emplace_back(proj.begin(), proj.begin() + cc + 1);
          if (ok) ++j;
        }
      }
      long cursor = known;
      while (cursor < j && iterations < o_.max_iterations && applied < rd_) {
        apply(Bc(L + cursor), Ac(applied));
        ++iterations;
        copy(w_.data(), Ac(applied));
        ++applied;
        const bool room = j < rd_;
        const bool ok = orthonormalize(L + j, L, room ? Bc(L + j) : nullptr, &proj);
        columns.emplace_back(proj.begin(), proj.begin() + cursor + 1);
        if (room && ok) ++j;
        ++cursor;
      }
      const int size = int(applied);  // the window without its last, unapplied vector
      if (size == 0) break;
      std::vector<cplx> P(size_t(size) * size, cplx(0, 0));
      for (int cc = 0; cc < int(columns.size()) && cc < size; ++cc)
        for (int r = 0; r <= cc && r < int(columns[size_t(cc)].size()); ++r) {
          const cplx v = c_ ? columns[size_t(cc)][size_t(r)] : cplx(columns[size_t(cc)][size_t(r)].real(), 0.0);
          P[size_t(r) * size + cc] = v;
          P[size_t(cc) * size + r] = std::conj(v);
        }
      std::vector<double> vals;
      std::vector<cplx> Uh;
      hermitian_eigen(size, P, vals, Uh);
      const long active = k - L;
      const long keep = std::max<long>(std::min<long>(active, rd_ / 2), 1);
      const auto idx = select_indices(vals, o_.target, int(std::min<long>(keep, size)), o_.shift);
      const int nsel = int(idx.size());
      // the selected Ritz coefficients (column-major size × nsel); a real
      // window takes real eigenvectors (phase of the largest entry removed)
      std::vector<double> Uc(size_t(size) * nsel * e_);
      for (int q = 0; q < nsel; ++q) {
        cplx ph(1, 0);
        if (!c_) {
          double best = 0;
          for (int r = 0; r < size; ++r)
            if (std::abs(Uh[size_t(r) * size + idx[q]]) > best) {
              best = std::abs(Uh[size_t(r) * size + idx[q]]);
              ph = Uh[size_t(r) * size + idx[q]] / best;
            }
        }
        for (int r = 0; r < size; ++r) {
          const cplx u = Uh[size_t(r) * size + idx[q]] / ph;
          if (c_) {
            Uc[2 * (size_t(q) * size + r)] = u.real();
            Uc[2 * (size_t(q) * size + r) + 1] = u.imag();
          } else {
            Uc[size_t(q) * size + r] = u.real();
          }
        }
      }
      QMBED_HIP(hipMemcpyAsync(U_.data(), Uc.data(), Uc.size() * 8, hipMemcpyHostToDevice, s_));
      lk::gemm_vc(c_, Bc(L), n_, size, U_.data(), nsel, Cv_.data(), s_);
      lk::gemm_vc(c_, Ac(0), n_, size, U_.data(), nsel, Ca_.data(), s_);
      for (int q = 0; q < nsel; ++q) {
        double* sq = sl_.data() + 4 * q;
        lk::nrm2sq_to_slot(c_, Cvc(q), n_, sq, ws_, s_);
        lk::normalize(c_, Cvc(q), sq, n_, Cvc(q), nullptr, nullptr, s_);
        lk::normalize(c_, Cac(q), sq, n_, Cac(q), nullptr, nullptr, s_);
        lk::redot_to_slot(c_, Cvc(q), Cac(q), n_, sq + 1, ws_, s_);
        lk::residual_to_slot(c_, Cac(q), Cvc(q), sq + 1, n_, sq + 2, ws_, s_);
      }
      QMBED_HIP(hipStreamSynchronize(s_));  // Uc is released below
      const double* cs = read(sl_.data(), size_t(4 * nsel));
      std::vector<double> value(nsel), residual(nsel);
      for (int q = 0; q < nsel; ++q) {
        const double rn = std::sqrt(cs[4 * q]);
        if (!std::isfinite(rn) || rn <= kEps) throw non_convergence(iterations, rn);
        value[q] = cs[4 * q + 1];
        residual[q] = cs[4 * q + 2];
      }
      std::vector<int> order(nsel);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return target_before(o_.target, value[a], value[b]); });
      last_residual = 0;
      for (long r = 0; r < active && r < nsel; ++r) last_residual = std::max(last_residual, residual[order[r]]);
      nret = 0;
      for (int r = 0; r < nsel; ++r) {
        const int q = order[r];
        const bool conv = residual[q] <= o_.tolerance * std::max(std::fabs(value[q]), 1.0);
        if (r < active && conv && L < k) {
          copy(w_.data(), Cvc(q));
          if (orthonormalize(L, L, Bc(L), nullptr)) {
            locked_value.push_back(value[q]);
            locked_residual.push_back(residual[q]);
            ++L;
          }
        } else if (nret < keep) {
          copy(Rc(nret), Cvc(q));
          copy(RAc(nret), Cac(q));
          ++nret;
        }
      }
      ++cycle;
    }
    if (L < k) throw non_convergence(iterations, last_residual);
    std::vector<int> order(static_cast<size_t>(L));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return target_before(o_.target, locked_value[a], locked_value[b]); });
    es_.iterations = iterations;
    vec_out.resize(size_t(k) * col);
    for (long q = 0; q < k; ++q) {
      es_.eigenvalues.push_back(locked_value[size_t(order[q])]);
      es_.residuals.push_back(locked_residual[size_t(order[q])]);
      copy(vec_out.data() + size_t(q) * col, Bc(order[q]));
    }
    es_.fp64_iterations = iterations;
    QMBED_HIP(hipStreamSynchronize(s_));
    return es_;
  }

 private:
  const Operator& op_;
  const EigshOptions& o_;
  const bool c_;
  const size_t n_, e_;
  hipStream_t s_;
  long rd_ = 0;
  lk::Workspace ws_;
  DeviceArray<double> B_, A_, R_, RA_, Cv_, Ca_, w_, h_, sl_, U_;
  PinnedArray<double> pin_;
  Eigensystem es_;

  double* Bc(long j) { return B_.data() + size_t(j) * e_ * n_; }
  double* Ac(long j) { return A_.data() + size_t(j) * e_ * n_; }
  double* Rc(long j) { return R_.data() + size_t(j) * e_ * n_; }
  double* RAc(long j) { return RA_.data() + size_t(j) * e_ * n_; }
  double* Cvc(long j) { return Cv_.data() + size_t(j) * e_ * n_; }
  double* Cac(long j) { return Ca_.data() + size_t(j) * e_ * n_; }
  void copy(double* dst, const double* src) {
    QMBED_HIP(hipMemcpyAsync(dst, src, e_ * n_ * 8, hipMemcpyDeviceToDevice, s_));
  }
  void apply(const double* x, double* y) {
    if (c_) op_.apply(reinterpret_cast<const cplx*>(x), reinterpret_cast<cplx*>(y), s_);
    else op_.apply(x, y, s_);
  }
  // count doubles of device memory, read back through pinned memory
  const double* read(const double* d, size_t count) {
    QMBED_HIP(hipMemcpyAsync(pin_.p, d, count * 8, hipMemcpyDeviceToHost, s_));
    QMBED_HIP(hipStreamSynchronize(s_));
    return pin_.p;
  }
  // Gram-Schmidt of w_ against B[0, cols) (classical, with the reference's
  // DGKS test for a second pass: src/solve/mod.rs:1519-1586), then w_/‖w_‖
  // into dest (skipped when dest is null). proj receives the first pass's
  // overlaps with B[first, cols). False when the remainder is dropped (not
  // finite, or ‖w‖ <= 1e-14).
  bool orthonormalize(long cols, long first, double* dest, std::vector<cplx>* proj) {
    double before, after;
    if (proj) proj->clear();
    if (cols == 0) {
      lk::nrm2sq_to_slot(c_, w_.data(), n_, sl_.data(), ws_, s_);
      before = after = std::sqrt(read(sl_.data(), 1)[0]);
      es_.reorthogonalization_passes += 1;
    } else {
      lk::lanczos_overlaps(c_, B_.data(), n_, int(cols), w_.data(), ws_, s_, true, nullptr, nullptr);
      lk::lanczos_subtract(c_, B_.data(), n_, int(cols), w_.data(), h_.data(), ws_, s_, true, nullptr, nullptr, true);
      lk::nrm2sq_to_slot(c_, w_.data(), n_, sl_.data(), ws_, s_);
      QMBED_HIP(hipMemcpyAsync(pin_.p, h_.data(), size_t(cols + 1) * e_ * 8, hipMemcpyDeviceToHost, s_));
      QMBED_HIP(hipMemcpyAsync(pin_.p + size_t(cols + 1) * e_, sl_.data(), 8, hipMemcpyDeviceToHost, s_));
      QMBED_HIP(hipStreamSynchronize(s_));
      const double* hh = pin_.p;
      before = std::sqrt(hh[size_t(cols) * e_]);
      after = std::sqrt(hh[size_t(cols + 1) * e_]);
      if (proj)
        for (long i = first; i < cols; ++i)
          proj->push_back(c_ ? cplx(hh[2 * i], hh[2 * i + 1]) : cplx(hh[i], 0.0));
      const bool second = before > kEps && after <= kDgks * before;
      es_.reorthogonalization_passes += second ? 2 : 1;
      if (second) {
        es_.conditional_second_passes += 1;
        lk::lanczos_overlaps(c_, B_.data(), n_, int(cols), w_.data(), ws_, s_, false, nullptr, nullptr);
        lk::lanczos_subtract(c_, B_.data(), n_, int(cols), w_.data(), h_.data(), ws_, s_, false, nullptr, nullptr,
                             true);
        lk::nrm2sq_to_slot(c_, w_.data(), n_, sl_.data(), ws_, s_);
        after = std::sqrt(read(sl_.data(), 1)[0]);
      }
    }
    if (!std::isfinite(after) || after <= 1e-14) return false;
    if (dest) lk::normalize(c_, w_.data(), sl_.data(), n_, dest, nullptr, nullptr, s_);
    return true;
  }
};

}  // namespace

// The thick restart from the given device start vectors (layout as the
// window: interleaved complex when c), and its dense fallback when the
// window is the whole space of at most 2048 rows within the budget.
Eigensystem thick_restart_eigsh(const Operator& op, const EigshOptions& o, const std::vector<const double*>& init,
                                bool c, DeviceArray<double>& vec_out, bool& complex_vectors, hipStream_t s) {
  const size_t n = op.dim();
  complex_vectors = c;
  try {
    ThickRestart tr(op, o, c, s);
    return tr.run(init, vec_out);
  } catch (Error& err) {
    if (err.kind != ErrorKind::NonConvergence) throw;
    if (n <= 2048 && size_t(o.max_iterations) >= n && (!o.krylov_dimension || size_t(*o.krylov_dimension) == n)) {
      Eigensystem es = dense_eigsh_route(op, o, vec_out, complex_vectors, s);
      es.iterations += err.iterations;
      return es;
    }
    throw;
  }
}

}  // namespace qmbed
