// Learned Lanczos warm start for parameter scans (scan.hip).
#pragma once

#include <future>

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
  // Network targets. Projections: the coefficients of each training ground
  // state over the snapshots. Continuation: the lowest eigenvector of the
  // family projected on the snapshots, Sᵀ H(c) S (eigenvector continuation),
  // at the training points and at `samples` seeded points of their bounding
  // box, so the network learns that map over the whole domain.
  // Correction: the step from the continuation coefficients to the training
  // ground state's own normalized projection on the snapshots,
  // Sᵀψ/‖Sᵀψ‖ − y(c), scaled to unit root-mean-square; the network's start
  // is then S·(y(c) + MLP(c)), the continuation start moved toward the best
  // start the snapshot space holds (the projection of the ground state).
  enum class Targets { Projections, Continuation, Correction };
  Targets targets = Targets::Projections;
  int samples = 512;
  // training points after the first `snapshots` start from the projection
  // on the snapshot space of those already solved
  bool provisional = true;
  // The snapshot space. Greedy: `snapshots` training ground states chosen by
  // pivoted Cholesky of their Gram matrix (largest residual first).
  // Principal: the `snapshots` leading principal components of all training
  // ground states (eigenvectors of the Gram matrix), the space that best
  // approximates them in the mean-square sense.
  enum class Space { Greedy, Principal };
  Space space = Space::Greedy;
};

class WarmStartScan {
 public:
  // `family` must stay alive and is re-evaluated by every call; s: the
  // stream all work runs on
  WarmStartScan(Operator& family, hipStream_t s, const ScanOptions& o);
  ~WarmStartScan();
  WarmStartScan(const WarmStartScan&) = delete;
  WarmStartScan& operator=(const WarmStartScan&) = delete;

  struct TrainReport {
    std::vector<double> energies;
    std::vector<long> iterations;
    std::vector<int> snapshot_indices;
    int snapshots = 0;
    int provisional_starts = 0;  // training solves started from the provisional space
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
  // How a point's Lanczos start is made.
  enum class Start {
    Plain,        // the deterministic start
    Network,      // v₀ = S·MLP(c)
    SnapshotRR,   // v₀ = S·y, y the lowest eigenvector of Sᵀ H(c) S
  };
  // ground state of one point from the chosen start; `y`: the point's
  // continuation coefficients when already computed (snapshot start)
  PointResult solve(const double* point, const EigshOptions& o, const EigshControls& ctl, Start start,
                    const std::vector<double>* y = nullptr);
  // ground states of n points: in lockstep batches (ctl.lockstep, any
  // start) where the family allows; otherwise in turn, and then with the
  // snapshot start the host eigensolve of point i + 1's projected family runs
  // on a helper thread while point i is solved on the device
  std::vector<PointResult> solve_many(const double* points, int n, const EigshOptions& o, const EigshControls& ctl,
                                      Start start);
  std::vector<PointResult> solve_lockstep(const double* points, int n, const EigshOptions& o, const EigshControls& ctl,
                                          Start start);
  // network outputs (snapshot coefficients) of n points
  std::vector<double> predict(const double* points, int n);

  bool trained() const { return trained_; }
  int snapshots() const { return ns_; }
  int parameters() const { return P_; }

 private:
  std::vector<cplx> coefficients_of(const double* point) const;
  // gauge-fixed lowest eigenvector of Sᵀ H(c) S (and, with gap, the
  // distance between its two lowest eigenvalues)
  std::vector<double> continuation(const double* point, double* gap = nullptr) const;
  // snapshot space of the first n training states (sets S_, ns_, proj_,
  // gauge_); Lm and sel receive the coefficients and the selected states
  void build_space(const double* psi, int n, std::vector<double>& Lm, std::vector<int>& sel);
  void principal_space(const double* psi, int n, DeviceArray<double>& dG, int want, std::vector<double>& Lm);
  void greedy_space(const double* psi, int n, DeviceArray<double>& dG, int want, std::vector<double>& Lm,
                    std::vector<int>& sel);
  void project_family();
  double gauge_sign(const double* d_psi);

  Operator& op_;
  hipStream_t s_;
  ScanOptions opt_;
  int P_ = 0, ns_ = 0;
  bool trained_ = false;
  DeviceArray<double> S_;      // N × ns orthonormal snapshot basis
  // Sᵀ H_part S (ns × ns, column-major) of the static part, then of each
  // component: Sᵀ H(c) S = proj_[0] + Σ_p c_p proj_[1 + p]
  std::vector<std::vector<double>> proj_;
  std::vector<double> gauge_;  // Sᵀ 1: the gauge sign of S·y is that of y·gauge_
  // per-solve device buffers, kept between solves
  DeviceArray<float> dx_, dy_;
  DeviceArray<double> dc_, v0_;
  PinnedArray<double> cpin_;  // staging of the start coefficients
  // lockstep batches: the points' tables, diagonals, starts and coefficients,
  // and the interleaved lanes of the windows' next vectors
  DeviceArray<double> lk_amps_, lk_diags_, lk_starts_, lk_y_, lk_lanes64_;
  DeviceArray<float> lk_lanes32_;
  DeviceArray<float> params_;  // network parameters
  std::vector<double> center_, scale_;
  double target_scale_ = 1.0;  // network outputs × target_scale_ = coefficients (or corrections)
  // the coefficients of the network's start at n points (host): the
  // network's outputs, plus the continuation under correction targets
  std::vector<std::vector<double>> network_coefficients(const double* points, int n);
  std::vector<std::vector<double>> continuations_of(const double* points, int n) const;
  // the gaps of Sᵀ H(c) S at n points (host threads)
  std::vector<double> gaps_of(const double* points, int n) const;
  PinnedArray<double> amp_pin_;  // staging of the points' amplitude tables
};

}  // namespace qmbed
