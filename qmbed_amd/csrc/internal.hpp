// Declarations shared between the library's translation units (not part of
// the public C ABI).
#pragma once

#include "qmbed.hpp"

namespace qmbed {

// flip groups holding the column entries c·a of every coupling
void compile_groups(const SpinBasisSpec& spec, const std::vector<OperatorSpec>& terms, const AssemblyChecks& checks,
                    bool nonempty, std::vector<FlipGroup>& diag, std::vector<FlipGroup>& off, bool& real);

// ---- vector kernels (vec.hip) ----
namespace dev {
// partial-sum workspace for reductions
struct Reducer {
  DeviceArray<double> partial;   // per-block partials
  DeviceArray<double> result;    // device result slots
  std::vector<double> host;
  explicit Reducer(size_t slots = 1024);
};

void fill(double* y, double v, size_t n, hipStream_t s);
void copy(double* y, const double* x, size_t n, hipStream_t s);
// y = a*x + y
void axpy(double* y, double a, const double* x, size_t n, hipStream_t s);
void scal(double* x, double a, size_t n, hipStream_t s);
// dot products (real); result to host (synchronizes)
double dot(const double* x, const double* y, size_t n, hipStream_t s, Reducer& r);
double nrm2(const double* x, size_t n, hipStream_t s, Reducer& r);
// h[0..m) = V^T w, V column-major (ld = n); result copied to host h (sync)
void gemv_t(const double* V, size_t n, int m, const double* w, double* d_h, hipStream_t s, Reducer& r);
// w -= V h (h on device), optionally returns ||w||^2 on device slot
void gemv_n_sub(const double* V, size_t n, int m, const double* d_h, double* w, hipStream_t s);
// y = V c  (c on device, m coefficients)
void gemv_n(const double* V, size_t n, int m, const double* d_c, double* y, hipStream_t s);
// FP64 -> FP32 and back
void to_f32(float* y, const double* x, size_t n, hipStream_t s);
void to_f64(double* y, const float* x, size_t n, hipStream_t s);
// complex helpers (interleaved)
void zaxpy(cplx* y, cplx a, const cplx* x, size_t n, hipStream_t s);
cplx zdot(const cplx* x, const cplx* y, size_t n, hipStream_t s, Reducer& r);  // conj(x)·y
double znrm2(const cplx* x, size_t n, hipStream_t s, Reducer& r);
void zscal(cplx* x, cplx a, size_t n, hipStream_t s);
void zdscal(cplx* x, double a, size_t n, hipStream_t s);
// complex versions of the multi-vector reductions (V complex, column-major)
void zgemv_t(const cplx* V, size_t n, int m, const cplx* w, cplx* d_h, hipStream_t s, Reducer& r);
void zgemv_n_sub(const cplx* V, size_t n, int m, const cplx* d_h, cplx* w, hipStream_t s);
void zgemv_n(const cplx* V, size_t n, int m, const cplx* d_c, cplx* y, hipStream_t s);
// out = V * C for real V (n×m) and complex coefficients C (m), complex out
void dzgemv_n(const double* V, size_t n, int m, const cplx* d_c, cplx* y, hipStream_t s);
// max |x_i| over complex vector (device -> host)
double zamax(const cplx* x, size_t n, hipStream_t s, Reducer& r);
// batched column norms / dots for nvec complex columns
void zdot_batch(const cplx* X, const cplx* Y, size_t n, int nvec, cplx* out_host, hipStream_t s, Reducer& r);
void znrm2_batch(const cplx* X, size_t n, int nvec, double* out_host, hipStream_t s, Reducer& r);
void zamax_batch(const cplx* X, size_t n, int nvec, double* out_host, hipStream_t s, Reducer& r);
}  // namespace dev

}  // namespace qmbed
