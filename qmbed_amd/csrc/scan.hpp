// Learned Lanczos warm start for parameter scans (scan.hip).
#pragma once

#include "qmbed.hpp"

namespace qmbed {

struct ScanOptions {
  int snapshots = 32;  // snapshot ground states (network outputs)
  int hidden1 = 64, hidden2 = 64;
  int epochs = 4000;  // full-batch Adam epochs
  float learning_rate = 2.0e-3f;
  // learning rate of the last epoch (exponential decay from learning_rate;
  // equal to it: constant)
  float final_learning_rate = 2.0e-3f;
  uint64_t seed = 0;  // network initialization
};

class WarmStartScan {
 public:
  // `family` must stay alive and is re-evaluated by every call; s: the
  // stream all work runs on
  WarmStartScan(Operator& family, hipStream_t s, const ScanOptions& o);

  struct TrainReport {
    std::vector<double> energies;
    std::vector<long> iterations;
    std::vector<int> snapshot_indices;
    int snapshots = 0;
    double solve_seconds = 0, gram_seconds = 0, train_seconds = 0;
    float loss_first = 0, loss_last = 0;
    double target_norm2 = 0;  // mean ‖target‖² (scale of the loss)
  };
  // points: n rows of one real coefficient per component
  TrainReport train(const double* points, int n, const EigshOptions& o, const EigshControls& ctl);

  struct PointResult {
    Eigensystem es;
    bool warm = false;      // started from the network's prediction
    bool fallback = false;  // prediction vanished: plain start used
    double seconds = 0;
  };
  // ground state of one point, from the learned start (warm) or the plain
  // deterministic start
  PointResult solve(const double* point, const EigshOptions& o, const EigshControls& ctl, bool warm);
  // network outputs (snapshot coefficients) of n points
  std::vector<double> predict(const double* points, int n);

  bool trained() const { return trained_; }
  int snapshots() const { return ns_; }
  int parameters() const { return P_; }

 private:
  std::vector<cplx> coefficients_of(const double* point) const;
  double gauge_sign(const double* d_psi);

  Operator& op_;
  hipStream_t s_;
  ScanOptions opt_;
  int P_ = 0, ns_ = 0;
  bool trained_ = false;
  DeviceArray<double> S_;      // N × ns orthonormal snapshot basis
  DeviceArray<float> params_;  // network parameters
  std::vector<double> center_, scale_;
};

}  // namespace qmbed
