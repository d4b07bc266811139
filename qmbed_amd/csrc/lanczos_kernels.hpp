// Device-scalar Lanczos primitives (lanczos_kernels.hip). `c` selects complex
// (interleaved double pairs) or real FP64 vectors.
#pragma once

#include "qmbed.hpp"

namespace qmbed {
namespace lk {

struct Workspace {
  DeviceArray<double> partial;
  DeviceArray<double> partial2;
  DeviceArray<double> slots;
  Workspace();
};

// slot = Re <x, y>
void redot_to_slot(bool c, const void* x, const void* y, size_t n, double* slot, Workspace& ws, hipStream_t s);
void nrm2sq_to_slot(bool c, const void* x, size_t n, double* slot, Workspace& ws, hipStream_t s);
// w -= a v + b u ; nrm2 = |w|^2
void three_term(bool c, void* w, const void* v, const void* u, const double* a_slot, const double* b_slot, size_t n,
                double* nrm2_slot, Workspace& ws, hipStream_t s);
// one classical Gram-Schmidt pass against V[:, 0..m): h = V^H w, w -= V h; nrm2 = |w|^2
void project_out(bool c, const void* V, size_t n, int m, void* w, void* d_h, double* nrm2_slot, Workspace& ws,
                 hipStream_t s);
// out = w / sqrt(nrm2); beta_slot = sqrt(nrm2); optional FP32 copy
void normalize(bool c, const void* w, const double* nrm2_slot, size_t n, void* out, float* out32, double* beta_slot,
               hipStream_t s);
// Y = V C (n×m · m×k)
void gemm_vc(bool c, const void* V, size_t n, int m, const void* d_C, int k, void* Y, hipStream_t s);
void to_f32(bool c, const void* x, size_t n, float* y, hipStream_t s);
void from_f32(bool c, const float* y, size_t n, void* x, hipStream_t s);
// slot = |y - lam x|
void residual_to_slot(bool c, const void* y, const void* x, const double* lam_slot, size_t n, double* slot,
                      Workspace& ws, hipStream_t s);

}  // namespace lk
}  // namespace qmbed
