// Error constructors (display strings follow QMBED src/error.rs:39-73) and
// small host utilities shared by the library.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>

#include "internal.hpp"

namespace qmbed {

Error invalid_operator(const std::string& s) {
  return Error(ErrorKind::InvalidOperator, "invalid or unregistered operator: " + s);
}
Error invalid_coupling(const std::string& s) { return Error(ErrorKind::InvalidCoupling, "invalid coupling: " + s); }
Error invalid_site(long site, long sites) {
  std::ostringstream os;
  os << "site " << site << " is outside a lattice with " << sites << " sites";
  return Error(ErrorKind::InvalidSite, os.str());
}
Error invalid_sector(const std::string& s) {
  return Error(ErrorKind::InvalidSector, "invalid Hilbert-space sector: " + s);
}
Error incompatible_symmetry(const std::string& s) {
  return Error(ErrorKind::IncompatibleSymmetry, "incompatible symmetry request: " + s);
}
Error non_hermitian() { return Error(ErrorKind::NonHermitian, "a Hermitian operator is required"); }
Error invalid_options(const std::string& s) { return Error(ErrorKind::InvalidOptions, "invalid options: " + s); }
std::string debug_string(const std::string& s) {
  std::string out = "\"";
  for (unsigned char ch : s) {
    switch (ch) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      case '\0': out += "\\0"; break;
      default:
        if (ch < 0x20 || ch == 0x7f) {
          char buf[16];
          std::snprintf(buf, sizeof buf, "\\u{%x}", unsigned(ch));
          out += buf;
        } else {
          out += char(ch);
        }
    }
  }
  return out + "\"";
}

Error dimension_mismatch(const std::string& s) {
  return Error(ErrorKind::DimensionMismatch, "dimension mismatch: " + s);
}
Error unsupported_backend(const std::string& s) {
  return Error(ErrorKind::UnsupportedBackend, "unsupported backend: " + s);
}
Error internal_state(const std::string& s) { return Error(ErrorKind::InternalState, "internal state error: " + s); }

// Rust's `{:e}` formatting of f64: shortest round-trip mantissa, exponent
// without padding or '+', e.g. 1.5e-7, 1e0, inf.
static std::string rust_exp(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  if (v == 0.0) return std::signbit(v) ? "-0e0" : "0e0";
  char buf[64];
  for (int prec = 0; prec < 18; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*e", prec, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::string s(buf);
  auto epos = s.find('e');
  std::string mant = s.substr(0, epos);
  int ex = std::atoi(s.c_str() + epos + 1);
  return mant + "e" + std::to_string(ex);
}

Error non_convergence(long iterations, double residual) {
  std::ostringstream os;
  os << "solver did not converge after " << iterations << " iterations; residual=" << rust_exp(residual);
  Error e(ErrorKind::NonConvergence, os.str());
  e.iterations = iterations;
  e.residual = residual;
  return e;
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::string msg = std::string("HIP call failed: ") + what + ": " + hipGetErrorString(e);
    throw Error(ErrorKind::InternalState, "internal state error: " + msg);
  }
}

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// ------------------------------------------------------------ launch timer --
namespace {

struct LaunchTimer {
  std::atomic<bool> on{false};
  std::mutex mu;
  std::vector<hipEvent_t> pool;  // events free for reuse
  struct Rec {
    std::string kernel;
    double bytes = 0;
    hipEvent_t a = nullptr, b = nullptr;
  };
  std::vector<Rec> recs;
  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    QMBED_HIP(hipEventCreate(&e));
    return e;
  }
  void recycle() {
    for (auto& r : recs) {
      if (r.a) pool.push_back(r.a);
      if (r.b) pool.push_back(r.b);
    }
    recs.clear();
  }
};

LaunchTimer& timer() {
  static LaunchTimer* t = new LaunchTimer;  // lives until the process ends
  return *t;
}

}  // namespace

int launch_timer_begin(const char* kernel, hipStream_t s, double bytes) {
  auto& t = timer();
  if (!t.on.load(std::memory_order_relaxed)) return -1;
  std::lock_guard<std::mutex> g(t.mu);
  LaunchTimer::Rec r;
  r.kernel = kernel;
  r.bytes = bytes;
  r.a = t.take();
  QMBED_HIP(hipEventRecord(r.a, s));
  t.recs.push_back(r);
  return int(t.recs.size()) - 1;
}

void launch_timer_end(int record, hipStream_t s) {
  if (record < 0) return;
  auto& t = timer();
  std::lock_guard<std::mutex> g(t.mu);
  if (size_t(record) >= t.recs.size()) return;
  auto& r = t.recs[size_t(record)];
  r.b = t.take();
  QMBED_HIP(hipEventRecord(r.b, s));
}

void launch_timer_switch(bool on) {
  auto& t = timer();
  std::lock_guard<std::mutex> g(t.mu);
  if (on) {
    for (auto& r : t.recs) {  // pending records of an earlier run complete first
      if (r.b) QMBED_HIP(hipEventSynchronize(r.b));
    }
    t.recycle();
  }
  t.on.store(on, std::memory_order_relaxed);
}

std::string launch_timer_report() {
  auto& t = timer();
  std::lock_guard<std::mutex> g(t.mu);
  struct Sum {
    long launches = 0;
    double ms = 0, bytes = 0;
  };
  std::map<std::string, Sum> sum;
  for (auto& r : t.recs) {
    if (!r.b) continue;
    QMBED_HIP(hipEventSynchronize(r.b));
    float ms = 0;
    QMBED_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    auto& e = sum[r.kernel];
    e.launches += 1;
    e.ms += double(ms);
    e.bytes += r.bytes;
  }
  std::ostringstream os;
  os.precision(17);
  os << '{';
  bool first = true;
  for (auto& kv : sum) {
    if (!first) os << ',';
    first = false;
    const Sum& e = kv.second;
    os << '"' << kv.first << "\":{\"launches\":" << e.launches << ",\"ms_total\":" << e.ms
       << ",\"ms_mean\":" << e.ms / double(e.launches) << ",\"bytes_total\":" << e.bytes
       << ",\"GBs\":" << (e.ms > 0 ? e.bytes / (e.ms * 1.0e6) : 0.0) << '}';
  }
  os << '}';
  return os.str();
}

}  // namespace qmbed
