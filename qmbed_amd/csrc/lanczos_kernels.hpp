// Device-scalar Lanczos primitives (lanczos_kernels.hip). `c` selects complex
// (interleaved double pairs) or real FP64 vectors.
#pragma once

#include "qmbed.hpp"

namespace qmbed {
namespace lk {

struct Workspace {
  DeviceArray<double> partial;   // overlap partials, one row of columns per row block
  DeviceArray<double> partial2;  // per-block scalar partials
  DeviceArray<double> partial3;  // per-block α_j partials of a Lanczos step
  DeviceArray<double> slots;
  Workspace();
};

// slot = Re <x, y> (its square root when root)
void redot_to_slot(bool c, const void* x, const void* y, size_t n, double* slot, Workspace& ws, hipStream_t s,
                   bool root = false);
void nrm2sq_to_slot(bool c, const void* x, size_t n, double* slot, Workspace& ws, hipStream_t s);
// out = w / sqrt(nrm2); beta_slot = sqrt(nrm2); optional FP32 copy; out may be
// null to write beta_slot only
void normalize(bool c, const void* w, const double* nrm2_slot, size_t n, void* out, float* out32, double* beta_slot,
               hipStream_t s);
// Y = V C (n×m · m×k)
void gemm_vc(bool c, const void* V, size_t n, int m, const void* d_C, int k, void* Y, hipStream_t s);
// C = Aᵀ B for real column-major A (n × p) and B (n × q), n ≫ p, q: the
// p × q result (column-major, device) from per-block partial products
// summed in a fixed order; `partial` is grown as needed.
void gemm_atb(const double* A, size_t n, int p, const double* B, int q, double* C, DeviceArray<double>& partial,
              hipStream_t s);
// The same product with its partials in a static device array of
// atb_scratch_capacity() doubles per device (no allocation). False, with
// nothing launched, when the shape needs more partials than that. Calls
// that overlap on one device race on the array.
bool gemm_atb_scratch(const double* A, size_t n, int p, const double* B, int q, double* C, hipStream_t s);
size_t atb_scratch_capacity();
void to_f32(bool c, const void* x, size_t n, float* y, hipStream_t s);
void from_f32(bool c, const float* y, size_t n, void* x, hipStream_t s, const int* gate = nullptr);
// slot = |y - lam x|
void residual_to_slot(bool c, const void* y, const void* x, const double* lam_slot, size_t n, double* slot,
                      Workspace& ws, hipStream_t s);

// One host-free Lanczos step j (m = j + 1 columns) follows the reference's
// order (QMBED src/solve/mod.rs:1030-1080): w = H·v_j; α_j = <v_j, w>
// (lanczos_alpha); w₀ = w − α_j v_j − β_{j−1} v_{j−1} is formed on the fly by
// lanczos_overlaps (overlaps with V and ‖w₀‖²) and lanczos_subtract
// (w ← w₀ − V h); lanczos_finish takes the DGKS decision against ‖w₀‖ and
// normalizes. The launches of a first pass take `gate` = the window's halt
// flag and are no-ops while it is set. When lanczos_finish finds that a
// second pass is needed it sets *halt = j + 1 instead of normalizing; the
// host then runs lanczos_overlaps/lanczos_subtract on w (no three-term part,
// no ‖w‖², ungated) and lanczos_finish(second_pass), which clears *halt, and
// relaunches the voided steps.

// α_j = Re <v_j, w> into the α slot (per-block partials, then one fixed-order
// sum)
void lanczos_alpha(bool c, const void* vj, const void* w, size_t n, Workspace& ws, hipStream_t s, const int* gate,
                   double* alpha);

// the three-term part of a first pass
struct ThreeTerm {
  const void* vj;
  const void* vjm1;         // null at j = 0
  const double* beta_prev;  // β_{j−1} slot (unused at j = 0)
  double* alpha;            // α_j slot, written by lanczos_alpha
};

// per-row-block overlaps of w₀ with V[:, 0..m) (and ‖w₀‖² when with_ww)
void lanczos_overlaps(bool c, const void* V, size_t n, int m, const void* w, Workspace& ws, hipStream_t s,
                      bool with_ww, const int* gate, const ThreeTerm* three);
// h = Σ partials (into d_h, m entries, + ‖w₀‖² at d_h[m] when with_ww;
// one reduction launch), then w ← w₀ − V h and per-block ‖w‖² partials
void lanczos_subtract(bool c, const void* V, size_t n, int m, void* w, void* d_h, Workspace& ws, hipStream_t s,
                      bool with_ww, const int* gate, const ThreeTerm* three, bool reduce = true);

// Lockstep windows (real vectors, first passes): one launch of a step phase
// for up to 16 windows, each window's blocks computing exactly what its own
// launch would. reduce = false above leaves the reduction of the overlaps to
// lanczos_hred_multi.
struct AlphaJob {
  const double* vj;
  const double* w;
  Workspace* ws;
  const int* gate;
  double* alpha;  // the step's α slot
};
void lanczos_alpha_multi(const AlphaJob* jobs, int nj, size_t n, hipStream_t s);
struct HredJob {
  Workspace* ws;
  int m;  // columns of the step's overlaps (‖w₀‖² follows at m)
  double* h;
  const int* gate;
};
void lanczos_hred_multi(const HredJob* jobs, int nj, size_t n, hipStream_t s);
struct FinishJob {
  const double* w;
  Workspace* ws;
  const double* h;
  int j;
  double* beta;
  int* halt;
  int* hist;
  double* out;
};
void lanczos_finish_multi(const FinishJob* jobs, int nj, size_t n, double eta, hipStream_t s);
// Starts of up to 16 lockstep windows in three launches: out = x / ‖x‖
// (each bit for bit as nrm2sq_to_slot and normalize give it), and
// zero[0..nzero) cleared (the window's flag area).
struct StartJob {
  const double* x;
  double* out;
  double* zero;
  int nzero;
};
void start_multi(const StartJob* jobs, int nj, size_t n, Workspace& ws, hipStream_t s);
// Up to 16 small copies in one launch (the read-backs of lockstep windows
// into host-visible pinned memory): dst[i] = src[i] for i < count.
struct CopyJob {
  const double* src;
  double* dst;
  int count;
};
void copy_multi(const CopyJob* jobs, int nj, hipStream_t s);
// The overlap pass (lanczos_overlaps with ‖w₀‖²) and the subtraction
// (lanczos_subtract, reduce = false) of the first pass of a step, for up to
// 16 lockstep windows in one launch each (the overlaps: one per column tile
// in use); every window's blocks compute exactly what its own launches would.
// m = j + 1 columns, at most one column chunk.
struct PassJob {
  const double* V;
  int m;
  double* w;
  Workspace* ws;
  const double* h;  // the reduced overlaps (subtraction)
  const int* gate;
  ThreeTerm three;
};
void lanczos_overlaps_multi(const PassJob* jobs, int nj, size_t n, hipStream_t s);
void lanczos_subtract_multi(const PassJob* jobs, int nj, size_t n, hipStream_t s);
// One lane of interleaved copies of a real vector: x64[i·stride + lane] (FP64)
// and x32[i·stride + lane] (FP32), either optional. Windows solved in
// lockstep keep their next vectors side by side this way, so that one
// gathered cache line serves every window's H·v.
struct Interleave {
  double* x64 = nullptr;
  float* x32 = nullptr;
  int stride = 1;
  int lane = 0;
};

// DGKS decision with threshold eta (first pass), β_j = ‖w‖, out = w / β_j
// (+ FP32 copy, + interleaved copies of a real vector)
void lanczos_finish(bool c, const void* w, size_t n, const void* d_h, int j, double eta, double* beta, int* halt,
                    int* hist, void* out, float* out32, Workspace& ws, hipStream_t s, bool second_pass,
                    const Interleave* il = nullptr);
// the interleaved copies of a real vector (one lane)
void scatter_lane(const double* x, size_t n, const Interleave& il, hipStream_t s);
// x[i·width + b] = src[b][i] for b < width (src[b] == nullptr: zero), as
// FP64 (x64) or FP32 (x32); width 1, 2, 4, 8 or 16
void interleave(const double* const* src, size_t n, int width, double* x64, float* x32, hipStream_t s);

}  // namespace lk
}  // namespace qmbed
