// GPU eigsh: dispatch (QMBED src/solve/mod.rs:2242-2452), fully
// reorthogonalized Lanczos windows (src/solve/mod.rs:1116-1456) and the dense
// small-problem route (src/solve/mod.rs:2198-2240) — all vector work on the
// device, only the m×m tridiagonal problem on the host.
//
// Differences from the reference loop, by design:
//  * reorthogonalization is classical Gram-Schmidt (two GEMV passes over V
//    per pass) instead of modified Gram-Schmidt; the DGKS second-pass
//    criterion ‖w‖ <= ‖w₀‖/√2 is the reference's;
//  * Ritz residual estimates β_j |s_{j,i}| are checked every few steps and
//    the window stops as soon as the requested pairs have converged (the
//    reference always fills its window, so `iterations` counts differ; the
//    returned pairs pass the same residual test ‖Hv − λv‖ <= tol·max(|λ|,1)).
//  * the FP32-matvec mode runs the Krylov window with single-precision H·v
//    and FP64 vectors/coefficients, then refines in FP64 from the FP32 Ritz
//    vectors (same acceptance test, FP64 Rayleigh quotient and residual).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>

#include "internal.hpp"
#include "lanczos_kernels.hpp"

namespace qmbed {

std::vector<double> tridiagonal_eigvec(const std::vector<double>& alpha, const std::vector<double>& beta, double lam);

static const double F64_EPS = 2.220446049250313e-16;
static const double DGKS = 0.7071067811865476;  // src/solve/mod.rs:57

void EigshOptions::validate(size_t dimension) const {
  if (eigenpairs <= 0 || size_t(eigenpairs) >= dimension)
    throw invalid_options("eigenpairs must be positive and smaller than the operator dimension");
  if (!std::isfinite(tolerance) || tolerance <= 0.0 || max_iterations <= 0)
    throw invalid_options("tolerance and max_iterations must be positive");
  if (krylov_dimension && (*krylov_dimension <= eigenpairs || size_t(*krylov_dimension) > dimension))
    throw invalid_options("krylov_dimension must exceed eigenpairs and not exceed dimension");
  if (target == SpectrumTarget::Shift && !std::isfinite(shift)) throw invalid_options("shift must be finite");
}

static int total_cmp(double a, double b) { return (a < b) ? -1 : (a > b ? 1 : 0); }

std::vector<int> select_indices(const std::vector<double>& v, SpectrumTarget target, int count, double shift) {
  const int n = int(v.size());
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  if (target == SpectrumTarget::BothEnds) {
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      int c = total_cmp(v[a], v[b]);
      return c != 0 ? c < 0 : a < b;
    });
    int lower = count / 2, upper = count - lower;
    std::vector<int> sel(idx.begin(), idx.begin() + std::min(lower, n));
    for (int i = 0; i < upper && i < n; ++i) sel.push_back(idx[n - 1 - i]);
    std::stable_sort(sel.begin(), sel.end(), [&](int a, int b) { return v[a] < v[b]; });
    return sel;
  }
  std::sort(idx.begin(), idx.end(), [&](int a, int b) {
    int c = 0;
    switch (target) {
      case SpectrumTarget::SmallestAlgebraic: c = total_cmp(v[a], v[b]); break;
      case SpectrumTarget::LargestAlgebraic: c = total_cmp(v[b], v[a]); break;
      case SpectrumTarget::SmallestMagnitude: c = total_cmp(std::fabs(v[a]), std::fabs(v[b])); break;
      case SpectrumTarget::LargestMagnitude: c = total_cmp(std::fabs(v[b]), std::fabs(v[a])); break;
      case SpectrumTarget::Shift: c = total_cmp(std::fabs(v[a] - shift), std::fabs(v[b] - shift)); break;
      default: break;
    }
    return c != 0 ? c < 0 : a < b;
  });
  idx.resize(std::min(count, n));
  return idx;
}

std::vector<double> selected_ritz_values(const std::vector<double>& a, const std::vector<double>& b,
                                         SpectrumTarget target, int k, double shift,
                                         std::vector<std::vector<double>>* vectors) {
  const int n = int(a.size());
  const bool low = target == SpectrumTarget::SmallestAlgebraic || target == SpectrumTarget::BothEnds ||
                   target == SpectrumTarget::LargestMagnitude;
  const bool high = target == SpectrumTarget::LargestAlgebraic || target == SpectrumTarget::BothEnds ||
                    target == SpectrumTarget::LargestMagnitude;
  std::vector<double> out;
  if ((low || high) && 2 * k < n) {
    // candidates: the k lowest and/or k highest eigenvalues, ascending
    std::vector<double> cand;
    if (low) cand = tridiagonal_eigenvalue_range(a, b, 0, k);
    if (high) {
      auto h = tridiagonal_eigenvalue_range(a, b, n - k, n);
      cand.insert(cand.end(), h.begin(), h.end());
    }
    for (int i : select_indices(cand, target, k, shift)) out.push_back(cand[size_t(i)]);
    if (!vectors) return out;
    bool close = false;
    for (size_t i = 0; i < out.size() && !close; ++i)
      for (size_t j = i + 1; j < out.size(); ++j)
        if (std::fabs(out[i] - out[j]) <= 1.0e-9 * std::max(std::fabs(out[i]), 1.0)) close = true;
    if (!close) {
      vectors->clear();
      for (double v : out) vectors->push_back(tridiagonal_eigvec(a, b, v));
      return out;
    }
    out.clear();
  }
  std::vector<double> vals, vecs;
  tridiagonal_eigen(a, b, vals, vectors ? &vecs : nullptr);
  const auto sel = select_indices(vals, target, k, shift);
  if (vectors) vectors->clear();
  for (int i : sel) {
    out.push_back(vals[size_t(i)]);
    if (vectors) {
      std::vector<double> v(static_cast<size_t>(n));
      for (int r = 0; r < n; ++r) v[size_t(r)] = vecs[size_t(r) * n + i];
      vectors->push_back(std::move(v));
    }
  }
  return out;
}

static bool target_before(SpectrumTarget t, double a, double b, double shift) {
  switch (t) {
    case SpectrumTarget::LargestAlgebraic: return b < a;
    case SpectrumTarget::SmallestMagnitude: return std::fabs(a) < std::fabs(b);
    case SpectrumTarget::LargestMagnitude: return std::fabs(b) < std::fabs(a);
    case SpectrumTarget::Shift: return std::fabs(a - shift) < std::fabs(b - shift);
    default: return a < b;
  }
}

namespace {

struct Candidate {
  double value;
  double residual;
  int column;  // column in the Ritz buffer
};

// Device and pinned host buffers of a Lanczos window, kept in a small
// per-device pool between solves so that a solve does not allocate (and free,
// which synchronizes the device) its Krylov basis every time. Buffers only
// grow; the pool keeps at most two sets per device.
struct WindowBuffers {
  int device = 0;
  DeviceArray<double> V, w, slots;
  DeviceArray<double> start, coef;  // eigsh start vector; Ritz coefficients
  DeviceArray<double> ritz;         // Ritz vectors of a convergence test
  DeviceArray<float> x32, y32;
  lk::Workspace ws;
  double* pin = nullptr;  // pinned read-back area
  size_t pin_n = 0;
  hipEvent_t ev = nullptr;
  double* pin2 = nullptr;  // pinned area of a convergence test (coefficients up, results down)
  size_t pin2_n = 0;
  hipEvent_t ev2 = nullptr;
  WindowBuffers() {
    QMBED_HIP(hipGetDevice(&device));
    QMBED_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    QMBED_HIP(hipEventCreateWithFlags(&ev2, hipEventDisableTiming));
  }
  ~WindowBuffers() {
    if (pin) (void)hipHostFree(pin);
    if (ev) (void)hipEventDestroy(ev);
    if (pin2) (void)hipHostFree(pin2);
    if (ev2) (void)hipEventDestroy(ev2);
  }
  // no copy through pin2 may be in flight
  void ensure_pin2(size_t n) {
    if (pin2_n >= n) return;
    if (pin2) QMBED_HIP(hipHostFree(pin2));
    pin2 = nullptr;
    pin2_n = 0;
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&pin2), n * sizeof(double), hipHostMallocDefault));
    pin2_n = n;
  }
  void ensure_pin(size_t n) {
    if (pin_n >= n) return;
    if (pin) QMBED_HIP(hipHostFree(pin));
    pin = nullptr;
    // coherent: a batched read-back kernel writes it directly
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&pin), n * sizeof(double), hipHostMallocCoherent));
    pin_n = n;
  }
};

template <typename T>
static void grow(DeviceArray<T>& a, size_t n) {
  if (a.size() < n) a.resize(n);
}

static std::mutex g_pool_mu;
static std::vector<std::unique_ptr<WindowBuffers>> g_pool;

static std::unique_ptr<WindowBuffers> acquire_buffers() {
  int dev = 0;
  QMBED_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_pool_mu);
  for (size_t i = 0; i < g_pool.size(); ++i) {
    if (g_pool[i]->device == dev) {
      auto b = std::move(g_pool[i]);
      g_pool.erase(g_pool.begin() + long(i));
      return b;
    }
  }
  return std::make_unique<WindowBuffers>();
}

// the caller has synchronized the stream that used b
static void release_buffers(std::unique_ptr<WindowBuffers> b) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  int same = 0;
  for (auto& e : g_pool) same += e->device == b->device ? 1 : 0;
  if (same < 20) g_pool.push_back(std::move(b));
}

// A pooled buffer set held for a scope; returned after the stream drains.
struct PooledBuffers {
  std::unique_ptr<WindowBuffers> b;
  hipStream_t s;
  explicit PooledBuffers(hipStream_t st) : b(acquire_buffers()), s(st) {}
  ~PooledBuffers() {
    if (hipStreamSynchronize(s) == hipSuccess) release_buffers(std::move(b));
  }
  PooledBuffers(const PooledBuffers&) = delete;
  PooledBuffers& operator=(const PooledBuffers&) = delete;
};

// Read-backs of lockstep windows in one launch: a window registers the copy
// of its slot area, flush() issues every registered copy with one kernel
// (into the windows' coherent pinned areas) and records one event. A window
// waits on the event of the flush that carried its copy; the events form a
// ring, and an event recorded again since is later in stream order, so
// waiting on it is still correct.
// The rings' events are kept per device between batches (creating 64
// events per batch cost a measurable share of a batch's host time).
static std::mutex g_ring_mu;
static std::vector<std::pair<int, hipEvent_t>> g_ring_events;

class PostBatch {
 public:
  explicit PostBatch(hipStream_t s) : s_(s) {
    QMBED_HIP(hipGetDevice(&device_));
    {
      std::lock_guard<std::mutex> lk(g_ring_mu);
      size_t k = 0;
      for (size_t i = 0; i < g_ring_events.size() && k < ev_.size();) {
        if (g_ring_events[i].first == device_) {
          ev_[k++] = g_ring_events[i].second;
          g_ring_events.erase(g_ring_events.begin() + long(i));
        } else {
          ++i;
        }
      }
    }
    for (auto& e : ev_)
      if (!e) QMBED_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  // the caller has synchronized the stream (no recorded event is pending)
  ~PostBatch() {
    std::lock_guard<std::mutex> lk(g_ring_mu);
    for (auto e : ev_)
      if (e) g_ring_events.push_back({device_, e});
  }
  PostBatch(const PostBatch&) = delete;
  PostBatch& operator=(const PostBatch&) = delete;
  // the generation of the flush that will carry this copy
  long add(const double* src, double* dst, size_t count) {
    jobs_.push_back({src, dst, int(count)});
    return gen_;
  }
  void flush() {
    if (jobs_.empty()) return;
    for (size_t i = 0; i < jobs_.size(); i += 16)
      lk::copy_multi(jobs_.data() + i, int(std::min<size_t>(16, jobs_.size() - i)), s_);
    QMBED_HIP(hipEventRecord(ev_[size_t(gen_ % kRing)], s_));
    jobs_.clear();
    ++gen_;
  }
  void wait(long g) {
    if (g >= gen_) flush();
    QMBED_HIP(hipEventSynchronize(ev_[size_t(g % kRing)]));
  }

 private:
  static constexpr long kRing = 64;
  hipStream_t s_;
  int device_ = 0;
  std::array<hipEvent_t, kRing> ev_{};
  std::vector<lk::CopyJob> jobs_;
  long gen_ = 0;
};

// One Lanczos window on the device. A step is six launches and no host
// round trip: H·v_j, the partials of α_j = <v_j, H·v_j>, the overlaps of
// w₀ = H·v_j − α_j v_j − β_{j−1} v_{j−1} with V (and ‖w₀‖²), their
// reduction, the subtraction w ← w₀ − V h, and a finishing kernel that takes
// the DGKS decision and normalizes w into v_{j+1}. Steps are launched
// speculatively on the assumption that no second pass is needed (the usual
// case); when one is, the finishing kernel raises a device halt flag that
// turns every later launch into a no-op, and the host, when it collects the
// next read-back, runs the second pass for that step and relaunches from
// there. Coefficients reach the host only through read-backs: `post` queues
// an asynchronous copy into pinned memory at the current step, `collect`
// waits for it (by then the device is already working on later steps) and
// parses the coefficients of the steps before it; breakdown (β ≤ 1e-14)
// truncates the window to the steps before it.
class LanczosWindow {
 public:
  // clear_flags = false: the caller clears flag_area() before the first step
  LanczosWindow(const Operator& op, bool complex, int m, bool fp32, hipStream_t s, double eta = DGKS,
                bool clear_flags = true)
      : op_(op), c_(complex), n_(op.dim()), m_(m), fp32_(fp32), eta_(eta), s_(s), buf_(acquire_buffers()) {
    const size_t esz = c_ ? 2 : 1;
    grow(buf_->V, esz * n_ * size_t(m_));
    grow(buf_->w, esz * n_);
    if (fp32_) {
      grow(buf_->x32, esz * n_);
      grow(buf_->y32, esz * n_);
    }
    // slots: [0,m) α, [m,2m) β, [2m, 2m+F) flags (m+1 ints: per-step
    // second-pass record, then halt), then h (2(m+1)), then scratch
    flag_doubles_ = (size_t(m_) + 2) / 2;
    const size_t head = 2 * size_t(m_) + flag_doubles_;
    grow(buf_->slots, head + 2 * size_t(m_ + 1) + 3 * size_t(m_ + 1) + 64);
    h_ = buf_->slots.data() + head;
    scratch_ = h_ + 2 * (m_ + 1);
    flags_ = reinterpret_cast<int*>(buf_->slots.data() + 2 * m_);
    if (clear_flags) QMBED_HIP(hipMemsetAsync(flags_, 0, flag_doubles_ * sizeof(double), s_));
    buf_->ensure_pin(head);
  }
  ~LanczosWindow() {
    if (hipStreamSynchronize(s_) == hipSuccess) release_buffers(std::move(buf_));
  }
  LanczosWindow(const LanczosWindow&) = delete;
  LanczosWindow& operator=(const LanczosWindow&) = delete;

  size_t esz() const { return c_ ? 2 : 1; }
  double* col(int j) const { return buf_->V.data() + esz() * n_ * size_t(j); }
  double* V() const { return buf_->V.data(); }
  double* w() const { return buf_->w.data(); }
  double* slots() const { return buf_->slots.data(); }
  double* flag_area() const { return reinterpret_cast<double*>(flags_); }
  size_t flag_doubles() const { return flag_doubles_; }
  lk::Workspace& workspace() { return buf_->ws; }

  // Install the start vector (device, FP64, already normalized) as column 0.
  // External matvec mode (windows solved in lockstep): the caller launches
  // H·v_j into w() before each step, reading the window's interleaved lane
  // of v_j, which this window keeps current (start and every normalization);
  // apply64 computes the FP64 products of the Rayleigh quotients.
  void set_external(const lk::Interleave& il, std::function<void(const double*, double*)> apply64) {
    if (c_) throw internal_state("lockstep windows are real");
    external_ = true;
    il_ = il;
    apply64_ = std::move(apply64);
  }
  bool external() const { return external_; }
  // a matvec launched by the caller for this window's next step
  void count_external(bool fp32) {
    ++matvecs_;
    if (fp32) ++matvecs32_;
  }
  const int* gate() const { return halt(); }

  void start_from(const double* d_v0) {
    QMBED_HIP(hipMemcpyAsync(col(0), d_v0, esz() * n_ * 8, hipMemcpyDeviceToDevice, s_));
    start_prepared();
  }
  // Make the window ready for a new start (lockstep slots refilled with the
  // next point): every per-solve counter and host record cleared; the
  // buffers stay. The window must have finished (no read-back, test or
  // step of the old start pending on the host side); its device work ahead
  // in the stream completes before anything of the new start.
  void reset_for_new_start() {
    launched_ = posted_ = size_ = 0;
    post_gen_ = 0;
    ritz_nc_ = ritz_nv_ = 0;
    passes_ = seconds_ = matvecs_ = matvecs64_ = matvecs32_ = 0;
    last_beta_ = 0;
    broke_ = false;
    alpha_.clear();
    beta_.clear();
  }
  // the caller has written the start into col(0) (and cleared the flags)
  void start_prepared() {
    if (external_ && (il_.x64 || il_.x32)) lk::scatter_lane(col(0), n_, il_, s_);
    use32_ = false;
    alpha_.clear();
    beta_.clear();
    launched_ = 0;
    size_ = 0;
  }

  // FP32 matvecs from the next step on (the caller decides when; see
  // run_window): the next column is converted once, and every later
  // normalization also writes the FP32 copy
  void set_fp32(bool on) {
    if (on == use32_) return;
    if (on) {
      if (!fp32_) throw internal_state("FP32 matvecs were not enabled for this window");
      if (launched_ < m_ && !external_) lk::to_f32(c_, col(launched_), n_, buf_->x32.data(), s_);
    }
    use32_ = on;
  }
  bool fp32_active() const { return use32_; }
  bool fp32_enabled() const { return fp32_; }

  void matvec(int j, const int* gate) {
    if (use32_) {
      float* x32 = buf_->x32.data();
      // FP32 input into FP64 accumulation and output where the operator has
      // that kernel; otherwise an FP32 product converted to FP64
      if (!op_.apply_mixed(x32, w(), c_, s_, gate)) {
        float* y32 = buf_->y32.data();
        if (c_) op_.apply(reinterpret_cast<const std::complex<float>*>(x32),
                          reinterpret_cast<std::complex<float>*>(y32), s_, gate);
        else op_.apply(x32, y32, s_, gate);
        lk::from_f32(c_, y32, n_, w(), s_, gate);
      }
      ++matvecs32_;
    } else {
      if (c_) op_.apply(reinterpret_cast<const cplx*>(col(j)), reinterpret_cast<cplx*>(w()), s_, gate);
      else op_.apply(col(j), w(), s_, gate);
    }
    ++matvecs_;
  }

  int* halt() const { return flags_ + m_; }

  // The phases of step j = launched() for windows stepping in lockstep (the
  // caller launched H·v_j into w(), and launches each phase of all windows
  // together): the same work as step().
  lk::ThreeTerm three_term() const {
    const int j = launched_;
    return lk::ThreeTerm{col(j), j > 0 ? col(j - 1) : nullptr, j > 0 ? slots() + m_ + (j - 1) : nullptr, slots() + j};
  }
  lk::AlphaJob alpha_job() { return {col(launched_), w(), &buf_->ws, halt(), slots() + launched_}; }
  lk::HredJob hred_job() { return {&buf_->ws, launched_ + 1, h_, halt()}; }
  lk::PassJob pass_job() { return {V(), launched_ + 1, w(), &buf_->ws, h_, halt(), three_term()}; }
  lk::FinishJob finish_job() {
    const int j = launched_;
    return {w(), &buf_->ws, h_, j, slots() + m_ + j, halt(), flags_ + j, j + 1 < m_ ? col(j + 1) : nullptr};
  }
  void advance() { ++launched_; }
  double eta() const { return eta_; }

  // Launch step j = launched() (requires launched() < m).
  void step() {
    const int j = launched_;
    const int* gate = halt();
    if (!external_) matvec(j, gate);
    const lk::ThreeTerm three{col(j), j > 0 ? col(j - 1) : nullptr, j > 0 ? slots() + m_ + (j - 1) : nullptr,
                              slots() + j};
    lk::lanczos_alpha(c_, col(j), w(), n_, buf_->ws, s_, gate, slots() + j);
    lk::lanczos_overlaps(c_, V(), n_, j + 1, w(), buf_->ws, s_, true, gate, &three);
    lk::lanczos_subtract(c_, V(), n_, j + 1, w(), h_, buf_->ws, s_, true, gate, &three);
    finish(j, false);
    ++launched_;
  }

  // Queue the read-back of the coefficients of the launched steps (with a
  // post batch: registered there, copied at its next flush, which comes
  // before this window launches again or waits for it).
  void post() {
    const size_t count = 2 * size_t(m_) + flag_doubles_;
    if (pb_) {
      post_gen_ = pb_->add(slots(), buf_->pin, count);
    } else {
      QMBED_HIP(hipMemcpyAsync(buf_->pin, slots(), count * 8, hipMemcpyDeviceToHost, s_));
      QMBED_HIP(hipEventRecord(buf_->ev, s_));
    }
    posted_ = launched_;
  }
  void set_post_batch(PostBatch* pb) { pb_ = pb; }
  int posted() const { return posted_; }

  // Wait for the posted read-back and parse it. A pending second pass (halt
  // flag set at step h) is queued here and the window rewinds to h + 1 (the
  // voided steps are relaunched by the caller). Otherwise `ended` is true
  // when the read-back covered the full window or a breakdown (the window is
  // then truncated to the step whose β fell to 1e-14 or below, as the
  // reference stops there).
  struct Sync {
    bool ended = false;
    bool rewound = false;
  };
  Sync collect() {
    if (pb_) pb_->wait(post_gen_);
    else QMBED_HIP(hipEventSynchronize(buf_->ev));
    const int L = posted_;
    const double* hb = buf_->pin;
    const int* hf = reinterpret_cast<const int*>(buf_->pin + 2 * m_);
    const int halted_at = hf[m_] - 1;  // -1: none
    if (halted_at >= L) throw internal_state("Lanczos halt flag beyond the launched steps");
    const int valid = halted_at >= 0 ? halted_at : L;  // steps with final α and β
    alpha_.clear();
    beta_.clear();
    Sync r;
    for (int j = 0; j < valid; ++j) {
      const double a = hb[size_t(j)], b = hb[size_t(m_ + j)];
      if (!std::isfinite(a) || !std::isfinite(b))
        throw internal_state("non-finite Lanczos coefficient at step " + std::to_string(j));
      if (b <= 1.0e-14) {
        for (int i = 0; i <= j; ++i) alpha_.push_back(hb[size_t(i)]);
        for (int i = 0; i < j; ++i) beta_.push_back(hb[size_t(m_ + i)]);
        last_beta_ = b;
        broke_ = true;
        size_ = j + 1;
        count_passes(hf, size_);
        r.ended = true;
        return r;
      }
    }
    if (halted_at >= 0) {
      // second pass for step h, then resume from h + 1
      const int h = halted_at;
      lk::lanczos_overlaps(c_, V(), n_, h + 1, w(), buf_->ws, s_, false, nullptr, nullptr);
      lk::lanczos_subtract(c_, V(), n_, h + 1, w(), h_, buf_->ws, s_, false, nullptr, nullptr);
      finish(h, true);
      launched_ = h + 1;
      r.rewound = true;
      return r;
    }
    for (int j = 0; j < L; ++j) {
      alpha_.push_back(hb[size_t(j)]);
      if (j + 1 < L) beta_.push_back(hb[size_t(m_ + j)]);
    }
    last_beta_ = hb[size_t(m_ + L - 1)];
    size_ = L;
    count_passes(hf, L);
    r.ended = L == m_;
    return r;
  }

  int launched() const { return launched_; }
  // completed steps known on the host (= columns spanned by T)
  int size() const { return size_; }
  const std::vector<double>& alpha() const { return alpha_; }
  const std::vector<double>& beta() const { return beta_; }
  double last_beta() const { return last_beta_; }
  long passes() const { return passes_; }
  long seconds() const { return seconds_; }
  long matvecs() const { return matvecs_; }
  long matvecs32() const { return matvecs32_; }
  bool broke() const { return broke_; }

  // largest residual estimate β|s_last| / max(|θ|, 1) of the selected pairs
  double max_estimate(int k, SpectrumTarget target, double shift) {
    const int sz = size_;
    if (sz <= k) return INFINITY;
    std::vector<double> a(alpha_.begin(), alpha_.begin() + sz);
    std::vector<double> b(beta_.begin(), beta_.begin() + std::max(0, sz - 1));
    double worst = 0;
    for (double val : selected_ritz_values(a, b, target, k, shift)) {
      auto v = tridiagonal_eigvec(a, b, val);
      const double est = last_beta_ * std::fabs(v[sz - 1]) / std::max(std::fabs(val), 1.0);
      worst = std::isfinite(est) ? std::max(worst, est) : INFINITY;
    }
    return worst;
  }

  // Ritz residual estimates for the selected pairs of the current T.
  bool estimates_converged(int k, SpectrumTarget target, double shift, double tol, double safety) {
    const int sz = size_;
    std::vector<double> a(alpha_.begin(), alpha_.begin() + sz);
    std::vector<double> b(beta_.begin(), beta_.begin() + std::max(0, sz - 1));
    const auto vals = selected_ritz_values(a, b, target, k, shift);
    if (int(vals.size()) < k) return false;
    for (double val : vals) {
      auto s = tridiagonal_eigvec(a, b, val);
      double est = last_beta_ * std::fabs(s[sz - 1]);
      if (!std::isfinite(est) || est > safety * tol * std::max(std::fabs(val), 1.0)) return false;
    }
    return true;
  }

  // Assemble Ritz vectors of the selected pairs (k columns, in the pooled
  // buffer ritz_vectors()), normalize, and compute FP64 Rayleigh quotients
  // and residuals.
  const double* ritz_vectors() const { return buf_->ritz.data(); }
  std::vector<Candidate> ritz(int k, SpectrumTarget target, double shift) {
    ritz_post(k, target, shift);
    return ritz_collect();
  }
  // ritz() in two halves: ritz_post launches the Ritz vectors and their FP64
  // tests, the results copied into pinned memory (the coefficients go up
  // through the same area); ritz_collect waits for them. Between the two the
  // window must not step (its w() holds the products of the test).
  void ritz_post(int k, SpectrumTarget target, double shift) { ritz_test_own(ritz_prepare(k, target, shift)); }
  // the tests of the nv prepared Ritz vectors with the window's own products
  void ritz_test_own(int nv) {
    rayleigh_launch(buf_->ritz.data(), nv);
    ritz_copy_results();
  }
  // the Ritz vectors of a test (ritz_post without their Rayleigh quotients);
  // returns their count
  int ritz_prepare(int k, SpectrumTarget target, double shift) {
    DeviceArray<double>& ritz_buf = buf_->ritz;
    const int sz = size_;
    std::vector<double> a(alpha_.begin(), alpha_.begin() + sz);
    std::vector<double> b(beta_.begin(), beta_.begin() + std::max(0, sz - 1));
    std::vector<std::vector<double>> vecs;
    const auto vals = selected_ritz_values(a, b, target, k, shift, &vecs);
    const size_t e = esz(), nv = vals.size(), nc = e * size_t(sz) * nv;
    buf_->ensure_pin2(nc + 3 * nv);  // the previous test has been collected
    double* C = buf_->pin2;
    std::fill(C, C + nc, 0.0);
    for (size_t q = 0; q < nv; ++q)
      for (int jj = 0; jj < sz; ++jj) C[e * (q * sz + jj)] = vecs[q][size_t(jj)];
    grow(buf_->coef, nc);
    QMBED_HIP(hipMemcpyAsync(buf_->coef.data(), C, nc * 8, hipMemcpyHostToDevice, s_));
    grow(ritz_buf, e * n_ * nv);
    lk::gemm_vc(c_, V(), n_, sz, buf_->coef.data(), int(nv), ritz_buf.data(), s_);
    ritz_nc_ = nc;
    ritz_nv_ = nv;
    return int(nv);
  }
  // the test's results into the pinned area, for ritz_collect
  void ritz_copy_results() {
    QMBED_HIP(hipMemcpyAsync(buf_->pin2 + ritz_nc_, scratch_, 3 * ritz_nv_ * sizeof(double), hipMemcpyDeviceToHost,
                             s_));
    QMBED_HIP(hipEventRecord(buf_->ev2, s_));
  }
  // A test of one Ritz vector whose product the caller launches (lockstep
  // windows: the tests of a tick share one H·v launch): verify_begin
  // normalizes the vector (returned), the caller writes H·y into w(), and
  // verify_end forms the Rayleigh quotient and residual and posts them.
  double* verify_begin() {
    double* y = buf_->ritz.data();
    rayleigh_pre(y, scratch_);
    return y;
  }
  void verify_end() {
    ++matvecs64_;
    rayleigh_post(buf_->ritz.data(), scratch_);
    ritz_copy_results();
  }
  std::vector<Candidate> ritz_collect() {
    QMBED_HIP(hipEventSynchronize(buf_->ev2));
    const double* h = buf_->pin2 + ritz_nc_;
    std::vector<Candidate> out;
    for (size_t q = 0; q < ritz_nv_; ++q) out.push_back({h[3 * q + 1], h[3 * q + 2], int(q)});
    return out;
  }

  // Normalize the nv columns of Y in place; FP64 Rayleigh quotients and
  // residual norms, read back with one synchronization.
  std::vector<std::pair<double, double>> rayleigh(double* Y, int nv) {
    rayleigh_launch(Y, nv);
    std::vector<double> host(3 * size_t(nv));
    QMBED_HIP(hipMemcpyAsync(host.data(), scratch_, host.size() * sizeof(double), hipMemcpyDeviceToHost, s_));
    QMBED_HIP(hipStreamSynchronize(s_));
    std::vector<std::pair<double, double>> out;
    for (int q = 0; q < nv; ++q) out.push_back({host[3 * size_t(q) + 1], host[3 * size_t(q) + 2]});
    return out;
  }
  // the launches of rayleigh(): scratch_[3q + 1] = Rayleigh quotient,
  // scratch_[3q + 2] = residual norm of column q
  void rayleigh_launch(double* Y, int nv) {
    if (3 * size_t(nv) > 3 * size_t(m_ + 1) + 64) throw internal_state("Rayleigh batch exceeds the slot area");
    const size_t e = esz();
    for (int q = 0; q < nv; ++q) {
      double* y = Y + e * n_ * size_t(q);
      double* sl = scratch_ + 3 * q;
      rayleigh_pre(y, sl);
      if (c_) op_.apply(reinterpret_cast<const cplx*>(y), reinterpret_cast<cplx*>(w()), s_);
      else if (apply64_) apply64_(y, w());
      else op_.apply(y, w(), s_);
      ++matvecs64_;
      rayleigh_post(y, sl);
    }
  }
  void rayleigh_pre(double* y, double* sl) {
    lk::nrm2sq_to_slot(c_, y, n_, sl, buf_->ws, s_);
    lk::normalize(c_, y, sl, n_, y, nullptr, nullptr, s_);
  }
  void rayleigh_post(const double* y, double* sl) {
    lk::redot_to_slot(c_, y, w(), n_, sl + 1, buf_->ws, s_);
    lk::residual_to_slot(c_, w(), y, sl + 1, n_, sl + 2, buf_->ws, s_);
  }

  long matvecs64() const { return matvecs64_; }

 private:
  void finish(int j, bool second_pass) {
    const bool next = j + 1 < m_;
    lk::lanczos_finish(c_, w(), n_, h_, j, eta_, slots() + m_ + j, halt(), flags_ + j, next ? col(j + 1) : nullptr,
                       (next && use32_ && !external_) ? buf_->x32.data() : nullptr, buf_->ws, s_, second_pass,
                       (next && external_) ? &il_ : nullptr);
  }
  void count_passes(const int* hf, int steps) {
    passes_ = 0;
    seconds_ = 0;
    for (int j = 0; j < steps; ++j) {
      passes_ += 1 + (hf[size_t(j)] ? 1 : 0);
      seconds_ += hf[size_t(j)] ? 1 : 0;
    }
  }

  const Operator& op_;
  bool c_;
  size_t n_;
  int m_;
  bool fp32_;
  double eta_;
  hipStream_t s_;
  std::unique_ptr<WindowBuffers> buf_;
  size_t flag_doubles_ = 0;
  int* flags_ = nullptr;  // [0,m) per-step second-pass record, m: halt (step + 1, 0 = none)
  double* h_ = nullptr;
  double* scratch_ = nullptr;
  std::vector<double> alpha_, beta_;
  double last_beta_ = 0;
  int launched_ = 0;
  int posted_ = 0;
  PostBatch* pb_ = nullptr;
  long post_gen_ = 0;
  size_t ritz_nc_ = 0, ritz_nv_ = 0;
  int size_ = 0;
  long passes_ = 0, seconds_ = 0, matvecs_ = 0, matvecs64_ = 0, matvecs32_ = 0;
  bool broke_ = false;
  bool use32_ = false;
  bool external_ = false;
  lk::Interleave il_;
  std::function<void(const double*, double*)> apply64_;
};

struct WindowResult {
  Eigensystem es;
  DeviceArray<double> vectors;  // k columns (device), complex if c
  bool ok = false;
  long iterations = 0;
  double fail_residual = 0;
  long fp32_matvecs = 0;  // of a window that failed after FP32 products
  // the start's ‖·‖ read at the window's first read-back was zero or not
  // finite (the window then stops there)
  bool bad_start = false;
  double start_norm = 0;
};

// One Lanczos window with in-window convergence checks, as a state machine:
// each tick launches the next step (when the window has room) and then
// posts, collects and acts on read-backs. run_window ticks one window to the
// end; windows solved in lockstep are ticked side by side after the caller
// has launched their shared matvec. A window throws NonConvergence (as
// Error) like the reference when the full window does not converge.
// start_norm2: a pinned host copy of the start's ‖·‖², queued before the
// window on the same stream and checked at its first read-back (nullptr:
// already checked).
// The convergence tests of the lockstep windows testing in one tick: their
// Ritz vectors' products in shared FP64 launches (as the steps' products,
// each lane computing what its own launch would), then each window's
// Rayleigh quotient and residual.
class VerifyBatch {
 public:
  VerifyBatch(const Operator& op, const std::vector<LockstepPoint>& pts, DeviceArray<double>& lanes64, hipStream_t s)
      : op_(op), pts_(pts), lanes64_(lanes64), s_(s) {}
  void add(int lane, LanczosWindow* w, const double* y) { items_.push_back({lane, w, y}); }
  void flush() {
    const size_t n = op_.dim();
    for (size_t g0 = 0; g0 < items_.size(); g0 += 8) {
      const int kg = int(std::min<size_t>(8, items_.size() - g0));
      const int width = kg <= 1 ? 1 : kg <= 2 ? 2 : kg <= 4 ? 4 : 8;
      const double* src[16] = {};
      const double* amps[16] = {};
      const double* diags[16] = {};
      double* ys[16] = {};
      for (int q = 0; q < kg; ++q) {
        const Item& it = items_[g0 + size_t(q)];
        src[q] = it.y;
        amps[q] = pts_[size_t(it.lane)].amps;
        diags[q] = pts_[size_t(it.lane)].diag;
        ys[q] = it.w->w();
      }
      const double* x = src[0];  // one lane: the vector itself
      if (width > 1) {
        lk::interleave(src, n, width, lanes64_.data(), nullptr, s_);
        x = lanes64_.data();
      }
      op_.apply_scan_batch(x, false, width, amps, diags, ys, nullptr, s_);
    }
    for (auto& it : items_) it.w->verify_end();
    items_.clear();
  }

 private:
  struct Item {
    int lane;
    LanczosWindow* w;
    const double* y;
  };
  const Operator& op_;
  const std::vector<LockstepPoint>& pts_;
  DeviceArray<double>& lanes64_;
  hipStream_t s_;
  std::vector<Item> items_;
};

class WindowRun {
 public:
  WindowRun(const Operator& op, bool c, const EigshOptions& o, int m, const double* d_start, bool fp32,
            const EigshControls& ctl, double accept_tol, WindowResult& out, hipStream_t s,
            const double* start_norm2 = nullptr, const lk::Interleave* il = nullptr,
            std::function<void(const double*, double*)> apply64 = {})
      : op_(op), c_(c), o_(o), m_(m), ctl_(ctl), accept_tol_(accept_tol), out_ptr_(&out), s_(s),
        start_norm2_(start_norm2), win_(op, c, m, fp32, s, ctl.dgks_eta, d_start != nullptr) {
    if (il) win_.set_external(*il, std::move(apply64));
    // d_start == nullptr: the caller writes the start into the window's
    // first column and clears its flags (lockstep windows, in one launch)
    if (d_start) win_.start_from(d_start);
    else win_.start_prepared();
    k_ = o.eigenpairs;
    next_check_ = std::max(k_ + 1, ctl.check_every);
    // Read-backs are posted at convergence checks, at the window's end, and
    // otherwise every `interval` steps (halved after a rewind, so frequent
    // second passes void few speculative steps). A posted read-back is
    // collected kLag steps later, when the device has long finished the copy,
    // so the device never waits for the host; a convergence found at a
    // collected point costs at most kLag extra steps.
    lag_ = ctl.read_lag > 0 ? ctl.read_lag : (op.dim() >= (size_t(1) << 20) ? 1 : 2);
    interval_ = std::max(1, ctl.check_every);
    next_sync_ = interval_;
  }
  bool done() const { return done_; }
  // A finished run takes a new start (the caller has written it into the
  // window's first column and cleared its flags); results go to `out` from
  // here on. Same checks, schedule and arithmetic as a fresh run.
  void restart(WindowResult& out) {
    out_ptr_ = &out;
    win_.reset_for_new_start();
    win_.start_prepared();
    next_check_ = std::max(k_ + 1, ctl_.check_every);
    interval_ = std::max(1, ctl_.check_every);
    next_sync_ = interval_;
    pending_ = pending_conv_ = done_ = false;
    verifying_ = verify_at_end_ = false;
    start_norm2_ = nullptr;
  }
  // the Ritz vectors of the window's test, in candidate order, into the
  // result (the window's own buffer is pooled)
  void copy_vectors(const std::vector<Candidate>& cands) {
    const size_t e = c_ ? 2 : 1;
    const size_t n = op_.dim();
    out().vectors.resize(e * n * cands.size());
    for (size_t q = 0; q < cands.size(); ++q)
      QMBED_HIP(hipMemcpyAsync(out().vectors.data() + e * n * q, win_.ritz_vectors() + e * n * cands[q].column,
                               e * n * 8, hipMemcpyDeviceToDevice, s_));
    QMBED_HIP(hipStreamSynchronize(s_));
  }
  // this tick launches a step
  bool stepping() const { return !done_ && !verifying_ && win_.launched() < m_; }
  // Convergence tests without a host wait (lockstep windows): a tick that
  // would test launches the test and returns; the window's next tick, when
  // the device has long run it (other windows' steps came after it), reads
  // the results and decides. The window launches nothing in between, so its
  // device work and results are those of the waiting form.
  void set_async_tests(bool on) { async_tests_ = on; }
  // the accepted Ritz vector (one wanted pair) is copied to d (device, e·N),
  // in stream order, instead of into the result
  void set_vector_destination(double* d) { vec_dst_ = d; }
  // tests of one real Ritz vector share the batch's product launches
  void set_verify_batch(VerifyBatch* vb, int lane) {
    vb_ = vb;
    vb_lane_ = lane;
  }
  LanczosWindow& window() { return win_; }

  // stepped: the caller has launched this tick's step (lockstep phases)
  void tick(bool stepped = false) {
    if (done_) return;
    if (verifying_) {
      verifying_ = false;
      decide(win_.ritz_collect(), verify_at_end_);
      return;
    }
    if (stepped) win_.advance();
    else if (win_.launched() < m_) win_.step();
    const int launched = win_.launched();
    const bool full = launched == m_;
    if (!pending_) {
      const bool conv = ctl_.early_stop && launched >= next_check_ && launched > k_;
      if (full || conv || launched >= next_sync_) {
        win_.post();
        pending_ = true;
        pending_conv_ = conv;
      }
    }
    if (!pending_ || (launched < win_.posted() + lag_ && !full)) return;
    pending_ = false;
    const auto st = win_.collect();
    if (start_norm2_) {
      const double nn = *start_norm2_;
      start_norm2_ = nullptr;
      if (!std::isfinite(nn) || std::sqrt(nn) <= F64_EPS) {
        out().bad_start = true;
        out().start_norm = std::sqrt(nn);
        done_ = true;
        return;
      }
    }
    if (st.rewound) {
      interval_ = std::max(1, interval_ / 2);
      next_sync_ = win_.launched() + interval_;
      return;
    }
    interval_ = std::min(std::max(1, ctl_.check_every), interval_ * 2);
    next_sync_ = win_.posted() + interval_;
    const bool at_end = st.ended;
    // FP32 matvecs once the wanted pairs have converged far enough that the
    // single-precision product errors, weighted by the small components of
    // the Ritz vectors on the remaining Krylov vectors, stay well below the
    // tolerance (relaxed inexact Krylov)
    if (win_.fp32_enabled() && !win_.fp32_active() && !at_end &&
        win_.max_estimate(k_, o_.target, o_.shift) <= ctl_.fp32_switch * accept_tol_)
      win_.set_fp32(true);
    if (!(at_end || pending_conv_)) return;
    const int sz = win_.size();
    if (!at_end) {
      if (!win_.estimates_converged(k_, o_.target, o_.shift, accept_tol_, 0.5)) {
        next_check_ = sz + ctl_.check_every;
        return;
      }
    }
    if (sz <= k_) {
      out().iterations = win_.matvecs();
      throw non_convergence(sz, INFINITY);
    }
    if (async_tests_) {
      const int nv = win_.ritz_prepare(k_, o_.target, o_.shift);
      if (vb_ && nv == 1 && !c_) vb_->add(vb_lane_, &win_, win_.verify_begin());
      else win_.ritz_test_own(nv);
      verifying_ = true;
      verify_at_end_ = at_end;
      return;
    }
    decide(win_.ritz(k_, o_.target, o_.shift), at_end);
  }

 private:
  // the FP64 test of the candidates: accept, continue, or fail
  void decide(std::vector<Candidate> cands, bool at_end) {
    const int sz = win_.size();
    std::stable_sort(cands.begin(), cands.end(), [&](const Candidate& a, const Candidate& b) {
      return target_before(o_.target == SpectrumTarget::BothEnds ? SpectrumTarget::SmallestAlgebraic : o_.target,
                           a.value, b.value, o_.shift);
    });
    double fail = 0;
    for (auto& cd : cands)
      if (cd.residual > accept_tol_ * std::max(std::fabs(cd.value), 1.0)) fail = std::max(fail, cd.residual);
    if (fail > 0 && !at_end) {
      // A Ritz pair whose estimate says converged but which fails the FP64
      // test after FP32 products: the single-precision errors already sit
      // in this window's tridiagonal matrix and more steps cannot remove
      // them, so the window ends here and the caller restarts in FP64 from
      // these Ritz vectors (accurate to about the FP32 floor).
      if (win_.matvecs32() > 0) {
        out().iterations = win_.matvecs();
        out().fp32_matvecs = win_.matvecs32();
        copy_vectors(cands);
        throw non_convergence(sz, fail);
      }
      next_check_ = sz + ctl_.check_every;
      return;
    }
    out().iterations = win_.matvecs();
    if (fail > 0) {
      out().fp32_matvecs = win_.matvecs32();
      copy_vectors(cands);
      Error e = non_convergence(sz, fail);
      throw e;
    }
    if (vec_dst_ && cands.size() == 1) {
      const size_t e = c_ ? 2 : 1, n = op_.dim();
      QMBED_HIP(hipMemcpyAsync(vec_dst_, win_.ritz_vectors() + e * n * cands[0].column, e * n * 8,
                               hipMemcpyDeviceToDevice, s_));
    } else if (ctl_.keep_vectors) {
      copy_vectors(cands);
    }
    for (auto& cd : cands) {
      out().es.eigenvalues.push_back(cd.value);
      out().es.residuals.push_back(cd.residual);
    }
    out().es.iterations = sz;
    out().es.reorthogonalization_passes = win_.passes();
    out().es.conditional_second_passes = win_.seconds();
    out().es.converged = true;
    out().es.fp32_iterations = win_.matvecs32();
    out().es.fp64_iterations = win_.matvecs() - win_.matvecs32() + win_.matvecs64();
    out().ok = true;
    done_ = true;
  }

 private:
  const Operator& op_;
  bool c_;
  const EigshOptions& o_;
  int m_;
  const EigshControls& ctl_;
  double accept_tol_;
  WindowResult* out_ptr_;
  hipStream_t s_;
  const double* start_norm2_;
  LanczosWindow win_;
  WindowResult& out() { return *out_ptr_; }
  int k_ = 1, next_check_ = 0, lag_ = 2, interval_ = 1, next_sync_ = 1;
  bool pending_ = false, pending_conv_ = false, done_ = false;
  bool async_tests_ = false, verifying_ = false, verify_at_end_ = false;
  double* vec_dst_ = nullptr;
  VerifyBatch* vb_ = nullptr;
  int vb_lane_ = 0;
};

void run_window(const Operator& op, bool c, const EigshOptions& o, int m, const double* d_start, bool fp32,
                const EigshControls& ctl, double accept_tol, WindowResult& out, hipStream_t s,
                const double* start_norm2 = nullptr) {
  WindowRun run(op, c, o, m, d_start, fp32, ctl, accept_tol, out, s, start_norm2);
  while (!run.done()) run.tick();
}

// Dense route for small problems: GPU materialization + rocSOLVER heevd.
Eigensystem dense_eigsh(const Operator& op, const EigshOptions& o, DeviceArray<double>& vec_out, bool& complex_out,
                        hipStream_t s, bool refine = true) {
  const size_t n = op.dim();
  DeviceArray<double> A(2 * n * n);
  op.materialize_dense(reinterpret_cast<cplx*>(A.data()), s);
  std::vector<cplx> host(n * n);
  QMBED_HIP(hipMemcpyAsync(host.data(), A.data(), n * n * 16, hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j)
      if (std::abs(host[i * n + j] - std::conj(host[j * n + i])) > 1.0e-12) throw non_hermitian();
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) throw internal_state("rocblas_create_handle failed");
  rocblas_set_stream(h, s);
  DeviceArray<double> D(n), E(n);
  DeviceArray<int> info(1);
  // row-major A is the column-major conjugate: eigenvectors come back
  // conjugated, eigenvalues unchanged
  // QR-based heev: the divide-and-conquer heevd returned non-finite
  // eigenpairs for n = 246 on gfx950 with this ROCm release
  rocblas_status st = rocsolver_zheev(h, rocblas_evect_original, rocblas_fill_upper, rocblas_int(n),
                                      reinterpret_cast<rocblas_double_complex*>(A.data()), rocblas_int(n), D.data(),
                                      E.data(), info.data());
  QMBED_HIP(hipStreamSynchronize(s));
  rocblas_destroy_handle(h);
  if (st != rocblas_status_success) throw internal_state("rocsolver_zheev failed");
  int hinfo = 0;
  QMBED_HIP(hipMemcpy(&hinfo, info.data(), sizeof(int), hipMemcpyDeviceToHost));
  if (hinfo != 0) throw non_convergence(0, INFINITY);
  std::vector<double> vals(n);
  QMBED_HIP(hipMemcpy(vals.data(), D.data(), n * 8, hipMemcpyDeviceToHost));
  std::vector<cplx> vecs(n * n);
  QMBED_HIP(hipMemcpy(vecs.data(), A.data(), n * n * 16, hipMemcpyDeviceToHost));
  auto sel = select_indices(vals, o.target, o.eigenpairs, o.shift);
  complex_out = !op.is_real();
  const size_t e = complex_out ? 2 : 1;
  vec_out.resize(e * n * sel.size());
  Eigensystem es;
  LanczosWindow rq(op, complex_out, 1, false, s);  // borrowed for Rayleigh refinement
  std::vector<double> col(e * n);
  for (size_t q = 0; q < sel.size(); ++q) {
    // column sel[q] of the column-major eigenvector matrix, conjugated back
    for (size_t i = 0; i < n; ++i) {
      cplx v = std::conj(vecs[size_t(sel[q]) * n + i]);
      if (complex_out) {
        col[2 * i] = v.real();
        col[2 * i + 1] = v.imag();
      } else {
        col[i] = v.real();
      }
    }
    double* y = vec_out.data() + e * n * q;
    QMBED_HIP(hipMemcpyAsync(y, col.data(), e * n * 8, hipMemcpyHostToDevice, s));
    auto r = rq.rayleigh(y, 1)[0];
    // eigh keeps the decomposition's eigenvalue and reports ‖Hv − λv‖ for it
    // (src/solve/mod.rs:586-605); eigsh refines it by the Rayleigh quotient
    const double lam = refine ? r.first : vals[size_t(sel[q])];
    es.eigenvalues.push_back(lam);
    es.residuals.push_back(refine ? r.second : std::sqrt(std::max(0.0, r.second * r.second +
                                                                           (r.first - lam) * (r.first - lam))));
  }
  es.iterations = 1;
  es.converged = true;
  return es;
}


}  // namespace

// eigh: the complete spectrum of the dense sector matrix (QMBED
// src/solve/mod.rs:586-620, the model's eigh at src/interop.rs:589-599): every
// eigenpair ascending, the residual ‖Hv − λv‖ of each, iterations 1 (0 for an
// empty operator), eigenvectors only when asked for.
Eigensystem eigh_all(const Operator& op, bool want_vectors, hipStream_t s) {
  Eigensystem es;
  const size_t n = op.dim();
  es.iterations = n ? 1 : 0;
  es.converged = true;
  if (n == 0) return es;
  if (n > 32768) throw unsupported_backend("dense eigh beyond 32768 rows is not built on the MI355X path");
  EigshOptions o;
  o.eigenpairs = int(n);
  o.target = SpectrumTarget::SmallestAlgebraic;
  DeviceArray<double> vecs;
  bool c = false;
  es = dense_eigsh(op, o, vecs, c, s, /*refine=*/false);
  es.iterations = 1;
  if (want_vectors) {
    const size_t e = c ? 2 : 1;
    std::vector<double> h(e * n * n);
    QMBED_HIP(hipMemcpy(h.data(), vecs.data(), h.size() * 8, hipMemcpyDeviceToHost));
    es.eigenvectors.assign(n, std::vector<cplx>(n));
    for (size_t q = 0; q < n; ++q)
      for (size_t i = 0; i < n; ++i)
        es.eigenvectors[q][i] = c ? cplx(h[e * n * q + 2 * i], h[e * n * q + 2 * i + 1]) : cplx(h[n * q + i], 0.0);
  }
  return es;
}

// the dense route, for the thick restart's fallback (thick_restart.cpp)
Eigensystem dense_eigsh_route(const Operator& op, const EigshOptions& o, DeviceArray<double>& vec_out,
                              bool& complex_out, hipStream_t s) {
  return dense_eigsh(op, o, vec_out, complex_out, s);
}

void eigsh_lockstep_stream(const Operator& op, const EigshOptions& o, const EigshControls& ctl, int slots,
                           const LockstepFeed& feed, DeviceArray<double>& lanes64, DeviceArray<float>& lanes32,
                           hipStream_t s) {
  const size_t n = op.dim();
  o.validate(n);
  if (slots < 1 || slots > 16) throw internal_state("a lockstep batch holds 1 to 16 windows");
  if (n <= 256 || o.target == SpectrumTarget::Shift) throw internal_state("lockstep windows need the Krylov route");
  const long m = o.krylov_dimension
                     ? std::min<long>(std::min<long>(*o.krylov_dimension, o.max_iterations), long(n))
                     : std::min<long>(std::min<long>(std::max<long>(8L * o.eigenpairs + 64, 256), long(n)),
                                      o.max_iterations);
  if (m <= o.eigenpairs) throw invalid_options("the effective Krylov dimension must exceed eigenpairs");
  const bool fp32 = ctl.precision == MatvecPrecision::FP32;
  const int max_width = op.scan_batch_max_width();
  if (lanes64.size() < 16 * n) lanes64.resize(16 * n);
  if (lanes32.size() < 16 * n) lanes32.resize(16 * n);
  // per slot: the point it holds (index, tables, start), its window and result
  std::vector<LockstepPoint> pts(static_cast<size_t>(slots));
  std::vector<int> point(static_cast<size_t>(slots), -1);
  std::vector<long> began(static_cast<size_t>(slots), 0);
  std::vector<WindowResult> res(static_cast<size_t>(slots));
  std::vector<std::unique_ptr<WindowRun>> runs(static_cast<size_t>(slots));
  std::vector<char> failed(static_cast<size_t>(slots), 0);
  std::vector<char> free_slot(static_cast<size_t>(slots), 0);  // finished, waiting for a refill
  const lk::Interleave no_lanes;
  PostBatch pb(s);
  VerifyBatch vb(op, pts, lanes64, s);
  // the starts of the slots (re)filled in one tick, normalized into their
  // windows' first columns and their flag areas cleared in three launches
  auto start_slots = [&](const std::vector<int>& fill) {
    std::vector<lk::StartJob> sj;
    for (int b : fill) {
      auto& w = runs[size_t(b)]->window();
      sj.push_back({pts[size_t(b)].start, w.col(0), w.flag_area(), int(w.flag_doubles())});
    }
    if (!sj.empty()) lk::start_multi(sj.data(), int(sj.size()), n, runs[size_t(fill[0])]->window().workspace(), s);
  };
  auto wire = [&](int b) {
    auto& r = runs[size_t(b)];
    r->window().set_post_batch(&pb);
    r->set_async_tests(true);
    r->set_verify_batch(&vb, b);
    r->set_vector_destination(pts[size_t(b)].vector);
  };
  {
    std::vector<int> fill;
    if (feed.prepare) feed.prepare(slots);
    for (int b = 0; b < slots; ++b) {
      point[size_t(b)] = feed.next(b, pts[size_t(b)]);
      if (point[size_t(b)] < 0) break;
      const LockstepPoint* pt = &pts[size_t(b)];
      auto apply64 = [&op, pt, s](const double* x, double* y) {
        const double* am[1] = {pt->amps};
        const double* dg[1] = {pt->diag};
        double* ys[1] = {y};
        op.apply_scan_batch(x, false, 1, am, dg, ys, nullptr, s);
      };
      runs[size_t(b)].reset(new WindowRun(op, false, o, int(m), nullptr, fp32, ctl, o.tolerance, res[size_t(b)], s,
                                          nullptr, &no_lanes, apply64));
      wire(b);
      fill.push_back(b);
    }
    start_slots(fill);
  }
  // every tick of a window launches a step or collects a read-back, so a
  // window ends within a few ticks per step of its window size
  const long max_ticks = 8 * (m + 8);
  for (long tick = 0;; ++tick) {
    // the windows stepping now, by the precision of their products (each
    // window's own FP32 switch, as in a single solve): their next vectors
    // interleaved per precision, and one launch over the index stream for
    // the FP32 lanes (up to 16) and up to 8 FP64 lanes together (an FP64
    // launch wider than 8 runs at half the occupancy; further FP64 lanes get
    // a launch of their own)
    bool active = false;
    struct Lane {
      const double* src;
      const double* amps;
      const double* diag;
      double* y;
      const int* gate;
    };
    std::vector<Lane> l32, l64;
    for (int b = 0; b < slots; ++b) {
      if (point[size_t(b)] < 0) continue;
      auto& r = *runs[size_t(b)];
      if (failed[size_t(b)] || r.done()) continue;
      active = true;
      if (tick - began[size_t(b)] > max_ticks) throw internal_state("lockstep Lanczos windows did not finish");
      if (!r.stepping()) continue;
      auto& w = r.window();
      const bool use32 = w.fp32_active();
      (use32 ? l32 : l64).push_back({w.col(w.launched()), pts[size_t(b)].amps, pts[size_t(b)].diag, w.w(), w.gate()});
      w.count_external(use32);
    }
    auto pow2 = [](int k) { return k <= 0 ? 0 : k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : k <= 8 ? 8 : 16; };
    auto launch = [&](const std::vector<Lane>& a32, size_t o32, int k32, const std::vector<Lane>& a64, size_t o64,
                      int k64, size_t lane64_base) {
      const int w32 = pow2(k32), w64 = pow2(k64);
      const double* src[16] = {};
      const double* amps[24] = {};
      const double* diags[24] = {};
      double* ys[24] = {};
      const int* gates[24] = {};
      for (int q = 0; q < k32; ++q) {
        const Lane& L = a32[o32 + size_t(q)];
        src[q] = L.src;
        amps[q] = L.amps;
        diags[q] = L.diag;
        ys[q] = L.y;
        gates[q] = L.gate;
      }
      if (w32) lk::interleave(src, n, w32, nullptr, lanes32.data(), s);
      const double* src64[16] = {};
      for (int q = 0; q < k64; ++q) {
        const Lane& L = a64[o64 + size_t(q)];
        src64[q] = L.src;
        amps[w32 + q] = L.amps;
        diags[w32 + q] = L.diag;
        ys[w32 + q] = L.y;
        gates[w32 + q] = L.gate;
      }
      double* x64 = lanes64.data() + lane64_base * n;
      if (w64) lk::interleave(src64, n, w64, x64, nullptr, s);
      op.apply_scan_mixed(lanes32.data(), w32, x64, w64, amps, diags, ys, gates, s);
    };
    {
      // chunks of at most max_width FP32 and 8 FP64 lanes, together when
      // their lane tables fit the LDS
      const size_t na = size_t(op.scan_batch_amp_count());
      auto fits = [&](int w) { return sizeof(double) * (size_t(w) * na + 81) <= 64 * 1024; };
      int o32 = 0, o64 = 0, k32 = int(l32.size()), k64 = int(l64.size());
      size_t base64 = 0;
      while (o32 < k32 || o64 < k64) {
        int c32 = std::min(k32 - o32, max_width), c64 = std::min(k64 - o64, 8);
        if (c32 > 0 && c64 > 0 && !fits(pow2(c32) + pow2(c64))) c64 = 0;
        launch(l32, size_t(o32), c32, l64, size_t(o64), c64, base64);
        o32 += c32;
        o64 += c64;
        if (c64 > 0) base64 = (base64 + 8) % 16;
      }
    }
    if (!active) break;
    // the rest of the step, each phase in one launch for every stepping
    // window: α, the overlaps (one launch per column tile in use), their
    // reductions, the subtractions, the finishing kernels
    std::vector<int> stepping;
    for (int b = 0; b < slots; ++b) {
      if (point[size_t(b)] < 0) continue;
      auto& r = *runs[size_t(b)];
      if (!failed[size_t(b)] && !r.done() && r.stepping()) stepping.push_back(b);
    }
    if (!stepping.empty()) {
      std::vector<lk::AlphaJob> aj;
      std::vector<lk::HredJob> hj;
      std::vector<lk::FinishJob> fj;
      for (int b : stepping) aj.push_back(runs[size_t(b)]->window().alpha_job());
      lk::lanczos_alpha_multi(aj.data(), int(aj.size()), n, s);
      std::vector<lk::PassJob> pj;
      for (int b : stepping) pj.push_back(runs[size_t(b)]->window().pass_job());
      lk::lanczos_overlaps_multi(pj.data(), int(pj.size()), n, s);
      for (int b : stepping) hj.push_back(runs[size_t(b)]->window().hred_job());
      lk::lanczos_hred_multi(hj.data(), int(hj.size()), n, s);
      lk::lanczos_subtract_multi(pj.data(), int(pj.size()), n, s);
      for (int b : stepping) fj.push_back(runs[size_t(b)]->window().finish_job());
      lk::lanczos_finish_multi(fj.data(), int(fj.size()), n, ctl.dgks_eta, s);
    }
    std::vector<char> stepped(static_cast<size_t>(slots), 0);
    for (int b : stepping) stepped[size_t(b)] = 1;
    for (int b = 0; b < slots; ++b) {
      if (point[size_t(b)] < 0) continue;
      auto& r = *runs[size_t(b)];
      if (failed[size_t(b)] || r.done()) continue;
      try {
        r.tick(stepped[size_t(b)] != 0);
      } catch (const Error&) {
        failed[size_t(b)] = 1;  // the caller solves this point on its own
      }
    }
    vb.flush();
    pb.flush();
    // finished windows hand their results over; free slots take the next
    // points once enough of them are free
    int running = 0;
    for (int b = 0; b < slots; ++b) {
      if (point[size_t(b)] < 0 || free_slot[size_t(b)]) continue;
      auto& r = *runs[size_t(b)];
      if (!failed[size_t(b)] && !r.done()) {
        ++running;
        continue;
      }
      LockstepResult out;
      out.ok = !failed[size_t(b)] && res[size_t(b)].ok;
      if (out.ok) out.es = res[size_t(b)].es;
      feed.done(point[size_t(b)], std::move(out));
      free_slot[size_t(b)] = 1;
    }
    int nfree = 0;
    for (int b = 0; b < slots; ++b) nfree += (point[size_t(b)] >= 0 && free_slot[size_t(b)]) ? 1 : 0;
    if (nfree > 0 && (nfree >= feed.group || running == 0)) {
      if (feed.prepare) feed.prepare(nfree);
      std::vector<int> fill;
      for (int b = 0; b < slots; ++b) {
        if (point[size_t(b)] < 0 || !free_slot[size_t(b)]) continue;
        point[size_t(b)] = feed.next(b, pts[size_t(b)]);
        if (point[size_t(b)] < 0) continue;
        auto& r = *runs[size_t(b)];
        free_slot[size_t(b)] = 0;
        res[size_t(b)] = WindowResult();
        failed[size_t(b)] = 0;
        began[size_t(b)] = tick + 1;
        r.restart(res[size_t(b)]);
        r.set_vector_destination(pts[size_t(b)].vector);
        fill.push_back(b);
      }
      start_slots(fill);
    }
    if (running == 0) {
      bool any = false;
      for (int b = 0; b < slots; ++b) any = any || (point[size_t(b)] >= 0 && !free_slot[size_t(b)]);
      if (!any) break;
    }
  }
  QMBED_HIP(hipStreamSynchronize(s));
}

std::vector<LockstepResult> eigsh_lockstep(const Operator& op, const EigshOptions& o, const EigshControls& ctl,
                                           const std::vector<LockstepPoint>& pts, DeviceArray<double>& lanes64,
                                           DeviceArray<float>& lanes32, hipStream_t s) {
  const int nb = int(pts.size());
  std::vector<LockstepResult> out(static_cast<size_t>(nb));
  int next = 0;
  LockstepFeed feed;
  feed.next = [&](int, LockstepPoint& pt) {
    if (next >= nb) return -1;
    pt = pts[size_t(next)];
    return next++;
  };
  feed.done = [&](int i, LockstepResult&& r) { out[size_t(i)] = std::move(r); };
  eigsh_lockstep_stream(op, o, ctl, nb, feed, lanes64, lanes32, s);
  return out;
}


// Balanced combination Σ_j e^{2πij/J}/√J v_j (src/solve/mod.rs:708-727),
// real part kept for real problems.
void balanced_start(const DeviceArray<double>& vecs, int k, size_t n, bool c, DeviceArray<double>& out,
                    hipStream_t s) {
  const size_t e = c ? 2 : 1;
  std::vector<double> coef(e * size_t(k), 0.0);
  for (int j = 0; j < k; ++j) {
    double ph = 6.283185307179586 * j / double(k);
    coef[e * j] = std::cos(ph) / std::sqrt(double(k));
    if (c) coef[e * j + 1] = std::sin(ph) / std::sqrt(double(k));
  }
  DeviceArray<double> dC(coef.size());
  QMBED_HIP(hipMemcpyAsync(dC.data(), coef.data(), coef.size() * 8, hipMemcpyHostToDevice, s));
  out.resize(e * n);
  lk::gemm_vc(c, vecs.data(), n, k, dC.data(), 1, out.data(), s);
  lk::Workspace ws;
  DeviceArray<double> sl(2);
  lk::nrm2sq_to_slot(c, out.data(), n, sl.data(), ws, s);
  lk::normalize(c, out.data(), sl.data(), n, out.data(), nullptr, nullptr, s);
  QMBED_HIP(hipStreamSynchronize(s));
}

// the smallest explicit window that runs the unrestarted tridiagonal
// Lanczos (TRIDIAGONAL_LANCZOS_WINDOW, src/solve/mod.rs)
constexpr long kTridiagonalWindow = 96;

Eigensystem shift_invert_eigsh(const Operator& op, const EigshOptions& o, const double* d_initial,
                               bool initial_complex, DeviceArray<double>& vec_out, bool& complex_vectors,
                               hipStream_t s);

Eigensystem thick_restart_eigsh(const Operator& op, const EigshOptions& o, const std::vector<const double*>& init,
                                bool c, DeviceArray<double>& vec_out, bool& complex_vectors, hipStream_t s);

Eigensystem eigsh_impl(const Operator& op, const EigshOptions& o, const double* d_initial, bool initial_complex,
                       const EigshControls& ctl, DeviceArray<double>& vec_out, bool& complex_vectors, hipStream_t s) {
  const size_t n = op.dim();
  o.validate(n);
  const bool dense = n <= 128 || (n <= 256 && !o.krylov_dimension);
  if (dense) return dense_eigsh(op, o, vec_out, complex_vectors, s);
  // the Shift target: Lanczos on (A − σI)⁻¹ (shift_invert.hip)
  if (o.target == SpectrumTarget::Shift)
    return shift_invert_eigsh(op, o, d_initial, initial_complex, vec_out, complex_vectors, s);
  const bool c = !op.is_real() || initial_complex;
  // explicit windows below the tridiagonal cut-off run the thick restart
  // (src/solve/mod.rs:2299-2327)
  if (o.krylov_dimension && *o.krylov_dimension < kTridiagonalWindow) {
    std::vector<const double*> init;
    if (d_initial) init.push_back(d_initial);
    return thick_restart_eigsh(op, o, init, c, vec_out, complex_vectors, s);
  }
  complex_vectors = c;
  const size_t e = c ? 2 : 1;
  // start vector (pooled scratch: no allocation per solve)
  PooledBuffers scratch(s);
  grow(scratch.b->start, e * n);
  grow(scratch.b->slots, 2);
  double* start = scratch.b->start.data();
  double* sl = scratch.b->slots.data();
  // the normalized start; its ‖·‖² is read back without a wait and checked
  // at the first window's first read-back (a zero or non-finite start then
  // fails as NonConvergence after 0 iterations)
  const double* raw = d_initial;
  if (!d_initial) {
    deterministic_start_device(start, n, o.seed, s, c);
    raw = start;
  }
  lk::nrm2sq_to_slot(c, raw, n, sl, scratch.b->ws, s);
  scratch.b->ensure_pin(1);
  const double* start_norm2 = scratch.b->pin;
  QMBED_HIP(hipMemcpyAsync(scratch.b->pin, sl, 8, hipMemcpyDeviceToHost, s));
  lk::normalize(c, raw, sl, n, start, nullptr, nullptr, s);

  const bool fp32 = ctl.precision == MatvecPrecision::FP32;
  // Windows: without krylov_dimension, min(max(8k+64, 256), N, budget),
  // doubled on non-convergence, dense fallback at the full space for
  // N <= 2048 (src/solve/mod.rs:2242-2328); with it, one window. A window
  // that fails after switching to FP32 products is repeated at the same size
  // in FP64 from the balanced combination of its Ritz vectors (the FP32
  // errors left a residual floor above the tolerance).
  DeviceArray<double> polish_start;
  auto run_expanding = [&](const double* d_start, bool use_fp32, double accept_tol, long budget) -> WindowResult {
    EigshOptions oo = o;
    const bool fixed = bool(o.krylov_dimension);
    long sub = fixed ? std::min<long>(std::min<long>(*o.krylov_dimension, o.max_iterations), long(n))
                     : std::min<long>(std::min<long>(std::max<long>(8L * o.eigenpairs + 64, 256), long(n)), budget);
    if (fixed && sub <= o.eigenpairs) throw invalid_options("the effective Krylov dimension must exceed eigenpairs");
    long spent = 0, fp32_spent = 0;
    for (;;) {
      WindowResult w2;
      try {
        run_window(op, c, oo, int(sub), d_start, use_fp32, ctl, accept_tol, w2, s, start_norm2);
        start_norm2 = nullptr;
        if (w2.bad_start) return w2;
        w2.es.iterations += spent;
        w2.es.fp32_iterations += fp32_spent;
        return w2;
      } catch (Error& err) {
        if (err.kind != ErrorKind::NonConvergence) throw;
        spent += err.iterations;
        if (use_fp32 && w2.fp32_matvecs > 0 && w2.vectors.size() > 0) {
          fp32_spent += w2.fp32_matvecs;
          balanced_start(w2.vectors, o.eigenpairs, n, c, polish_start, s);
          d_start = polish_start.data();
          use_fp32 = false;
          // the FP64 repeat stays within the iteration budget, for a fixed
          // window as for an expanding one
          const long remaining = budget > spent ? budget - spent : 0;
          sub = std::min(sub, remaining);
          if (sub <= o.eigenpairs) throw non_convergence(spent, err.residual);
          continue;
        }
        if (fixed) throw;
        if (sub == long(n) && n <= 2048) {
          WindowResult w3;
          bool cv = false;
          w3.es = dense_eigsh(op, o, w3.vectors, cv, s);
          w3.es.iterations += spent;
          w3.ok = true;
          return w3;
        }
        long remaining = budget > spent ? budget - spent : 0;
        long next = std::min(std::min(sub * 2, long(n)), remaining);
        if (next <= sub) throw non_convergence(spent, err.residual);
        sub = next;
      }
    }
  };

  WindowResult wr = run_expanding(start, fp32, o.tolerance, o.max_iterations);
  if (wr.bad_start) throw non_convergence(0, wr.start_norm);
  vec_out = std::move(wr.vectors);
  return wr.es;
}

Eigensystem eigsh(const Operator& op, const EigshOptions& o, const cplx* initial, const EigshControls& ctl,
                  bool want_vectors, hipStream_t s) {
  const size_t n = op.dim();
  DeviceArray<double> d_init;
  bool init_complex = false;
  if (initial) {
    for (size_t i = 0; i < n; ++i)
      if (!std::isfinite(initial[i].real()) || !std::isfinite(initial[i].imag()))
        throw invalid_options("eigsh initial vector must contain only finite values");
    double nn = 0;
    for (size_t i = 0; i < n; ++i) nn += std::norm(initial[i]);
    if (!std::isfinite(std::sqrt(nn)) || std::sqrt(nn) <= F64_EPS)
      throw invalid_options("eigsh initial vector must have nonzero finite norm");
    for (size_t i = 0; i < n; ++i)
      if (std::fabs(initial[i].imag()) > 1.0e-14) init_complex = true;
    const bool c = !op.is_real() || init_complex;
    std::vector<double> h(c ? 2 * n : n);
    for (size_t i = 0; i < n; ++i) {
      if (c) {
        h[2 * i] = initial[i].real();
        h[2 * i + 1] = initial[i].imag();
      } else {
        h[i] = initial[i].real();
      }
    }
    d_init.resize(h.size());
    QMBED_HIP(hipMemcpyAsync(d_init.data(), h.data(), h.size() * 8, hipMemcpyHostToDevice, s));
  }
  DeviceArray<double> vecs;
  bool cv = false;
  Eigensystem es = eigsh_impl(op, o, initial ? d_init.data() : nullptr, init_complex, ctl, vecs, cv, s);
  if (want_vectors) {
    const size_t e = cv ? 2 : 1;
    std::vector<double> h(e * n * es.eigenvalues.size());
    QMBED_HIP(hipMemcpy(h.data(), vecs.data(), h.size() * 8, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < es.eigenvalues.size(); ++q) {
      std::vector<cplx> v(n);
      for (size_t i = 0; i < n; ++i) v[i] = cv ? cplx(h[e * (q * n + i)], h[e * (q * n + i) + 1]) : cplx(h[q * n + i], 0);
      es.eigenvectors.push_back(std::move(v));
    }
  }
  return es;
}

void validate_eigsh_initial(const std::vector<cplx>& v, size_t dimension) {
  if (v.size() != dimension) throw dimension_mismatch("eigsh initial vector does not match the operator");
  double nn = 0;
  for (const cplx& z : v) {
    if (!std::isfinite(z.real()) || !std::isfinite(z.imag()))
      throw invalid_options("eigsh initial vector must contain only finite values");
    nn += std::norm(z);
  }
  if (!std::isfinite(std::sqrt(nn)) || std::sqrt(nn) <= F64_EPS)
    throw invalid_options("eigsh initial vector must have nonzero finite norm");
}

std::vector<cplx> balanced_subspace_start(const std::vector<std::vector<cplx>>& sub) {
  if (sub.empty()) throw invalid_options("initial subspace must not be empty");
  const size_t n = sub[0].size();
  const double count = double(sub.size());
  std::vector<cplx> out(n, cplx(0, 0));
  for (size_t j = 0; j < sub.size(); ++j) {
    const cplx coef = std::polar(1.0 / std::sqrt(count), 2.0 * M_PI * double(j) / count);
    for (size_t i = 0; i < n; ++i) out[i] += coef * sub[j][i];
  }
  auto normalize = [](std::vector<cplx>& v) {
    double nn = 0;
    for (const cplx& z : v) nn += std::norm(z);
    const double nr = std::sqrt(nn);
    if (!std::isfinite(nr) || nr <= F64_EPS) return false;
    for (cplx& z : v) z /= nr;
    return true;
  };
  if (!normalize(out)) {
    out = sub[0];
    if (!normalize(out)) throw non_convergence(0, 0.0);
  }
  return out;
}

Eigensystem eigsh_subspace(const Operator& op, const EigshOptions& o, const std::vector<std::vector<cplx>>& sub,
                           const EigshControls& ctl, bool want_vectors, hipStream_t s) {
  const size_t n = op.dim();
  o.validate(n);
  if (sub.empty()) return eigsh(op, o, nullptr, ctl, want_vectors, s);
  for (const auto& v : sub) validate_eigsh_initial(v, n);
  if (n <= 128 || (n <= 256 && !o.krylov_dimension)) return eigsh(op, o, nullptr, ctl, want_vectors, s);
  if (sub.size() == 1) return eigsh(op, o, sub[0].data(), ctl, want_vectors, s);
  if (o.target == SpectrumTarget::Shift)
    throw unsupported_backend("shift-invert targets from several start vectors (the thick restart) are not built on "
                              "the MI355X path");
  // several vectors: the thick restart keeps them as a subspace
  bool c = !op.is_real();
  for (const auto& v : sub)
    for (const cplx& z : v)
      if (std::fabs(z.imag()) > 1.0e-14) c = true;
  const size_t e = c ? 2 : 1;
  DeviceArray<double> d(e * n * sub.size());
  std::vector<double> h(e * n);
  std::vector<const double*> init;
  for (size_t j = 0; j < sub.size(); ++j) {
    for (size_t i = 0; i < n; ++i) {
      if (c) {
        h[2 * i] = sub[j][i].real();
        h[2 * i + 1] = sub[j][i].imag();
      } else {
        h[i] = sub[j][i].real();
      }
    }
    QMBED_HIP(hipMemcpy(d.data() + j * e * n, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    init.push_back(d.data() + j * e * n);
  }
  DeviceArray<double> vecs;
  bool cv = false;
  Eigensystem es = thick_restart_eigsh(op, o, init, c, vecs, cv, s);
  if (want_vectors) {
    const size_t ev = cv ? 2 : 1;
    std::vector<double> hv(ev * n * es.eigenvalues.size());
    QMBED_HIP(hipMemcpy(hv.data(), vecs.data(), hv.size() * 8, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < es.eigenvalues.size(); ++q) {
      std::vector<cplx> v(n);
      for (size_t i = 0; i < n; ++i)
        v[i] = cv ? cplx(hv[ev * (q * n + i)], hv[ev * (q * n + i) + 1]) : cplx(hv[q * n + i], 0);
      es.eigenvectors.push_back(std::move(v));
    }
  }
  return es;
}

void EigshWorkspace::set(size_t dim, std::vector<std::vector<cplx>> vectors) {
  for (const auto& v : vectors) validate_eigsh_initial(v, dim);
  dimension = dim;
  subspace = std::move(vectors);
}

Eigensystem eigsh_workspace(const Operator& op, const EigshOptions& o, EigshWorkspace& ws, const EigshControls& ctl,
                            hipStream_t s) {
  const size_t n = op.dim();
  Eigensystem es;
  if (ws.dimension == n && !ws.subspace.empty()) {
    if ((!o.krylov_dimension || *o.krylov_dimension >= kTridiagonalWindow) && ws.subspace.size() > 1) {
      const auto start = balanced_subspace_start(ws.subspace);
      o.validate(n);
      validate_eigsh_initial(start, n);
      es = eigsh(op, o, start.data(), ctl, true, s);
    } else {
      es = eigsh_subspace(op, o, ws.subspace, ctl, true, s);
    }
  } else {
    es = eigsh(op, o, nullptr, ctl, true, s);
  }
  ws.dimension = n;
  ws.subspace = es.eigenvectors;
  return es;
}

Eigensystem eigsh_on_device(const Operator& op, const EigshOptions& o, const double* d_initial, bool initial_complex,
                            const EigshControls& ctl, DeviceArray<double>& vectors, bool& complex_vectors,
                            hipStream_t s) {
  return eigsh_impl(op, o, d_initial, initial_complex, ctl, vectors, complex_vectors, s);
}

Eigensystem eigsh_device(const Operator& op, const EigshOptions& o, const double* d_initial, const EigshControls& ctl,
                         double* d_vectors_out, hipStream_t s) {
  DeviceArray<double> vecs;
  bool cv = false;
  Eigensystem es = eigsh_impl(op, o, d_initial, false, ctl, vecs, cv, s);
  if (d_vectors_out) {
    const size_t e = cv ? 2 : 1;
    QMBED_HIP(hipMemcpyAsync(d_vectors_out, vecs.data(), e * op.dim() * es.eigenvalues.size() * 8,
                             hipMemcpyDeviceToDevice, s));
    QMBED_HIP(hipStreamSynchronize(s));
  }
  return es;
}

}  // namespace qmbed
