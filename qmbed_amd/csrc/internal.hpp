// Declarations shared between the library's translation units (not part of
// the public C ABI).
#pragma once

#include "qmbed.hpp"

namespace qmbed {

// flip groups holding the column entries c·a of every coupling
void compile_groups(const SpinBasisSpec& spec, const std::vector<OperatorSpec>& terms, const AssemblyChecks& checks,
                    bool nonempty, std::vector<FlipGroup>& diag, std::vector<FlipGroup>& off, bool& real);
// operator family: parts[0] = static terms, parts[1+p] = component p; every
// group carries one table per part (FlipGroup::parts) and amp = Σ parts
void compile_family_groups(const SpinBasisSpec& spec, const std::vector<std::vector<OperatorSpec>>& parts,
                           const AssemblyChecks& checks, bool nonempty, std::vector<FlipGroup>& diag,
                           std::vector<FlipGroup>& off, bool& real);

// ---- sector-batched action (hv.hip) ----
// One launch applies many Hermitian sector operators, each to its own segment
// of concatenated complex FP64 vectors (sector i owns rows [offsets[i],
// offsets[i] + dim_i)), over the sectors' compiled index streams. The
// operators must stay alive; all must share one symmetry kind (translation,
// parity, both or none).
class SectorBatchAction {
 public:
  static constexpr int kRows = 256;  // rows per block
  SectorBatchAction(const std::vector<Operator*>& ops, const std::vector<size_t>& offsets, hipStream_t s);
  int sectors() const { return count_; }
  // block count of one sector (kRows rows each)
  uint32_t blocks_of(int sector) const { return blocks_[size_t(sector)]; }
  // y = H x over the listed blocks: d_items[b] = sector << 16 | block
  void apply(const uint32_t* d_items, int nitems, const double* x, double* y, hipStream_t s) const;

 private:
  int count_ = 0;
  int sym_ = 0;
  size_t smem_ = 0;
  std::vector<uint32_t> blocks_;
  DeviceArray<uint8_t> sec_;  // per-sector kernel arguments
};

// ---- launch timer (common.cpp) ----
// While switched on, the instrumented launches (the batched scan H·v and
// the lockstep windows' overlap and subtraction passes) are
// bracketed by two HIP events on their own stream; the report resolves the
// pairs and sums their elapsed times per kernel instance. Off, a launch pays
// one relaxed atomic load.
// Returns a record index, or -1 when the timer is off. `bytes`: the
// launch's algorithmic bytes (0: not stated), summed per kernel instance.
int launch_timer_begin(const char* kernel, hipStream_t s, double bytes = 0.0);
void launch_timer_end(int record, hipStream_t s);
void launch_timer_switch(bool on);
// {"kernel": {"launches": n, "ms_total": t, "ms_mean": t / n, "bytes_total":
// b, "GBs": b / t}, ...}; synchronizes the recorded events
std::string launch_timer_report();

// ---- FP32 MLP of the learned warm start (mlp.hip) ----
int mlp_parameter_count(int in, int h1, int h2, int out);
// Y [n][out] = MLP(X [n][in]) on the device
void mlp_forward_device(int in, int h1, int h2, int out, const float* d_params, const float* d_x, int n, float* d_y,
                        hipStream_t s);
// epochs > 0: full-batch Adam on (X, T) with the learning rate multiplied
// by lr_decay after every epoch, parameters updated in place, the loss of
// every epoch into d_loss_hist (optional); epochs == 0: one gradient into
// d_grad and the loss into d_loss
void mlp_train_device(int in, int h1, int h2, int out, float* d_params, const float* d_x, const float* d_t, int n,
                      int epochs, float lr, float lr_decay, float* d_grad, float* d_loss, float* d_loss_hist,
                      hipStream_t s);

}  // namespace qmbed
