// Core host-side types of the MI355X spin-chain engine.
//
// Mirrors the reference's SpinBasis1D / OperatorSpec / OperatorBuilder /
// EigshOptions / ExpmActionPlan surface (QMBED src/basis/mod.rs:1018-1590,
// src/operator/mod.rs:18-319 and 1941-2216, src/solve/mod.rs:14-526) with
// device-resident state: every basis, operator and vector lives in HBM and is
// acted on by the gfx950 kernels in the *.hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <functional>
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace qmbed {

using cplx = std::complex<double>;

// ------------------------------------------------------------------ errors --
// Display strings follow QmbedError (QMBED src/error.rs:39-73); the C ABI
// returns them verbatim.
enum class ErrorKind {
  InvalidOperator,
  InvalidCoupling,
  InvalidSite,
  StateNotInBasis,
  InvalidSector,
  IncompatibleSymmetry,
  NonHermitian,
  InvalidOptions,
  DimensionMismatch,
  UnsupportedBackend,
  InternalState,
  NonConvergence,
  Archive,
};

struct Error : std::exception {
  ErrorKind kind;
  std::string message;
  long iterations = 0;
  double residual = 0.0;
  Error(ErrorKind k, std::string m) : kind(k), message(std::move(m)) {}
  const char* what() const noexcept override { return message.c_str(); }
};

Error invalid_operator(const std::string& s);
Error invalid_coupling(const std::string& s);
Error invalid_site(long site, long sites);
Error invalid_sector(const std::string& s);
Error incompatible_symmetry(const std::string& s);
Error non_hermitian();
Error invalid_options(const std::string& s);
Error dimension_mismatch(const std::string& s);
// a string as Rust's `{:?}` prints it (quoted, with escapes) for error texts
std::string debug_string(const std::string& s);
Error unsupported_backend(const std::string& s);
Error internal_state(const std::string& s);
Error non_convergence(long iterations, double residual);

void hip_check(hipError_t e, const char* what);
#define QMBED_HIP(call) ::qmbed::hip_check((call), #call)

// ------------------------------------------------------ device buffers -----
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  ~DeviceArray() { release(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) QMBED_HIP(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Page-locked host buffer (asynchronous copies), grown on demand.
template <typename T>
struct PinnedArray {
  T* p = nullptr;
  size_t n = 0;
  PinnedArray() = default;
  PinnedArray(const PinnedArray&) = delete;
  PinnedArray& operator=(const PinnedArray&) = delete;
  void ensure(size_t want) {
    if (n >= want) return;
    if (p) QMBED_HIP(hipHostFree(p));
    p = nullptr;
    n = 0;
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(want, 1) * sizeof(T),
                            hipHostMallocDefault));
    n = want;
  }
  ~PinnedArray() {
    if (p) (void)hipHostFree(p);
  }
};

// ------------------------------------------------------------------ basis --
enum class Normalization { AngularMomentum = 0, Pauli = 1, PauliCartesian = 2 };

// One generator of a general symmetry sector (the C ABI's SymmetryRequest,
// QMBED bindings/capi/src/lib.rs:906-911; LatticeSymmetryMap,
// src/basis/mod.rs:2967-3060): site `source` moves to destinations[source]
// and its digit d becomes local_permutations[source][d] (empty: identity on
// every site); `sector` is the character index q of e^{-2πi q / period}.
struct SymmetryMapSpec {
  std::vector<long> destinations;
  std::vector<std::vector<long>> local_permutations;
  int sector = 0;
};

struct SpinBasisSpec {
  int sites = 0;
  int spin_twice = 1;
  std::optional<int> up;
  std::vector<int> up_sectors;  // union of up sectors (sorted, unique); empty = none
  std::optional<int> momentum;
  std::optional<int> parity;
  Normalization normalization = Normalization::AngularMomentum;
  // general symmetry sector (GeneralBasis + SymmetryReducer); cannot be
  // combined with momentum or parity
  std::vector<SymmetryMapSpec> symmetries;
  // the QuSpin row convention (bindings/capi/src/lib.rs:3412-3422): rows in
  // descending order of the representatives, and (general sectors) the
  // maximum state of each orbit as its representative
  bool reverse = false;
};

// Symmetry-sector data consumed by the kernels (all integers; characters are
// tabulated on the host).
struct SectorGeometry {
  int L = 0;
  uint64_t mask = 0;
  bool wide = false;  // 64-bit states (32 <= L <= 63)
  int kmom = -1;     // canonical momentum, -1 when no translation symmetry
  int parity = 0;    // 0 none, +1/-1
  int n_div = 0;     // proper divisors of L in ascending order (period search)
  int divisors[16] = {0};
  // general symmetry sector: the group generated by the maps, as n_group
  // elements s -> P(s) XOR flip (element 0 is the identity); n_chars
  // distinct character values; representatives are orbit maxima when
  // max_rep (minima otherwise)
  int n_group = 0;
  int n_chars = 1;
  bool max_rep = false;
  // rows in descending order of the representatives
  bool reversed = false;
};

// SpinBasis1D: QMBED src/basis/mod.rs:1018-1035, 1312-1412.
// Representatives are the minimum integers of their orbits, sorted ascending
// (src/basis/mod.rs:785-847); they and the per-row orbit sizes live in HBM.
class SpinBasis1D {
 public:
  static std::shared_ptr<SpinBasis1D> build(const SpinBasisSpec& spec, hipStream_t stream = nullptr);

  size_t len() const { return n_; }
  const SpinBasisSpec& spec() const { return spec_; }
  const SectorGeometry& geometry() const { return geo_; }
  bool symmetric() const { return geo_.kmom >= 0 || geo_.parity != 0 || geo_.n_group > 0; }
  bool general() const { return geo_.n_group > 0; }
  bool reversed() const { return geo_.reversed; }
  // character phases are real when the momentum is 0 or L/2
  bool real_characters() const {
    if (general()) return real_chars_;
    return geo_.kmom < 0 || geo_.kmom == 0 || 2 * geo_.kmom == geo_.L;
  }
  const uint32_t* states() const { return states_.data(); }
  // wide sectors (L >= 32): the representatives as 64-bit words, ascending
  // and in row order; states() and the search tables are empty then
  bool wide() const { return geo_.wide; }
  const uint64_t* states64() const { return states64_.data(); }
  const uint64_t* row_states64() const { return reversed() ? rstates64_.data() : states64_.data(); }
  const uint8_t* orbits() const { return orbits_.data(); }
  // the representatives and orbit sizes in row order (the ascending tables
  // above, or their reversal when reversed())
  const uint32_t* row_states() const { return reversed() ? rstates_.data() : states_.data(); }
  const uint8_t* row_orbits() const { return reversed() ? rorbits_.data() : orbits_.data(); }
  // general sectors: per group element the site destinations (64 bytes),
  // the destination-side flip mask and the character-table index of
  // χ(g⁻¹); the distinct character values (index 0 holds 1)
  const uint8_t* group_dest() const { return gdest_.data(); }
  const uint64_t* group_flip() const { return gflip_.data(); }
  const uint8_t* group_phase() const { return gphase_.data(); }
  // per element k or (k | 0x40) when every element moves the sites as a
  // rotation j -> j + k or a reflection of the ring (the element is then the
  // state, or its bit reversal, rotated left by k), else null
  const uint8_t* group_rot() const { return grot_.size() ? grot_.data() : nullptr; }
  const std::vector<cplx>& characters() const { return chars_; }
  std::vector<uint32_t> states_host() const;
  std::vector<uint8_t> orbits_host() const;
  // row order (descending when reversed())
  std::vector<uint32_t> row_states_host() const;
  // row order as 64-bit words, for sectors of either width
  std::vector<uint64_t> row_states_host64() const;
  std::vector<uint8_t> row_orbits_host() const;

  // State directory of a 32-bit sector (the H·v kernels' lookup): the top
  // min(L, kDirCellBits) bits of a value name its cell; cell c splits its
  // value range into 2^(cb − s_c) sub-buckets of 2^s_c values, s_c the
  // largest shift with at most kDirBucket states in every sub-bucket.
  // cells() holds per cell (first slot, s_c); directory() one aligned
  // 16-byte slot per sub-bucket: the index of its first state and its
  // states (ascending, padded with all ones), so one 16-byte read answers a
  // lookup.
  static constexpr int kDirCellBits = 11;
  static constexpr int kDirBucket = 3;
  const uint32_t* directory_cells() const { return dir_cells_.data(); }
  const uint32_t* directory() const { return dir_.data(); }
  int directory_shift() const { return dir_shift_; }  // cb = L − cell bits

 private:
  SpinBasisSpec spec_;
  SectorGeometry geo_;
  size_t n_ = 0;
  DeviceArray<uint32_t> states_;
  DeviceArray<uint8_t> orbits_;
  DeviceArray<uint32_t> rstates_;
  DeviceArray<uint8_t> rorbits_;
  DeviceArray<uint64_t> states64_, rstates64_;
  DeviceArray<uint8_t> gdest_;
  DeviceArray<uint64_t> gflip_;
  DeviceArray<uint8_t> gphase_;
  DeviceArray<uint8_t> grot_;
  std::vector<cplx> chars_;
  bool real_chars_ = true;
  DeviceArray<uint32_t> dir_cells_, dir_;
  int dir_shift_ = 0;
  static void build_directory(SpinBasis1D& b, hipStream_t stream);
};

// --------------------------------------------------------------- operator --
struct Coupling {
  cplx coefficient;
  std::vector<int> sites;
};

// OperatorSpec: QMBED src/operator/mod.rs:204-260 (symbols are the compact
// one-character spelling 'I' 'n' 'z' '+' '-' 'x' 'y').
struct OperatorSpec {
  std::string symbols;
  std::vector<Coupling> couplings;
};

// A named, parameterized part of an operator family: H(c) = H_static +
// Σ_p c_p H_p (QMBED bindings/capi/src/lib.rs:783-787 TermComponentRequest,
// src/operator/mod.rs:2663-2690 QuantumComponent). Without a default the
// component's coefficient is required at every evaluation.
struct ComponentSpec {
  std::string name;
  std::vector<OperatorSpec> terms;
  bool has_default = false;
  cplx default_value{0.0, 0.0};
};

struct AssemblyChecks {
  bool hermiticity = true;
  bool particle_conservation = true;
  bool symmetry_compatibility = true;
};

enum class MatrixFormat { Dense, Csc, Csr, Dia, MatrixFree };

// One flip group: every coupling whose spin-1/2 action flips the same set of
// bits. The gather-form amplitude for row state r is amp[pattern(r)], where
// pattern packs the bits of r at `sites` (bit i <- site sites[i]).
struct FlipGroup {
  uint64_t flip = 0;
  int nsites = 0;
  int sites[5] = {0, 0, 0, 0, 0};
  std::vector<cplx> amp;  // 1 << nsites entries
  // operator families: the table of every part (static first, then one per
  // component) with unit coefficients; amp = Σ_part coefficient · parts[part]
  std::vector<std::vector<cplx>> parts;
};

// Matrix-free Hermitian operator on one spin sector, applied on the fly by
// the gfx950 H·v kernels (no stored matrix).
class Operator {
 public:
  size_t dim() const { return basis_->len(); }
  bool is_real() const { return real_; }
  const SpinBasis1D& basis() const { return *basis_; }
  std::shared_ptr<const SpinBasis1D> basis_ptr() const { return basis_; }
  const std::vector<FlipGroup>& diagonal_groups() const { return diag_; }
  const std::vector<FlipGroup>& offdiagonal_groups() const { return off_; }
  MatrixFormat format() const { return format_; }

  // y = H x on `stream`. Pointers are device pointers. Complex vectors are
  // interleaved (re, im) pairs. FP32 variants compute in single precision.
  // `gate` (optional device int): a nonzero value at launch time makes the
  // action a no-op (used by speculative Lanczos steps)
  void apply(const double* x, double* y, hipStream_t s, const int* gate = nullptr) const;  // real (requires is_real)
  void apply(const cplx* x, cplx* y, hipStream_t s, const int* gate = nullptr) const;      // complex vectors
  void apply(const float* x, float* y, hipStream_t s, const int* gate = nullptr) const;    // real FP32
  void apply(const std::complex<float>* x, std::complex<float>* y, hipStream_t s, const int* gate = nullptr) const;
  // Mixed precision: FP32 input (real, or interleaved complex when
  // complex_vectors), FP64 accumulation and output. Only the compiled-stream
  // mode has this kernel: returns false (nothing launched) otherwise.
  bool apply_mixed(const float* x, double* y, bool complex_vectors, hipStream_t s, const int* gate = nullptr) const;
  // nvec columns (stride = dim), complex; one lookup shared across columns
  void apply_batch(const cplx* x, cplx* y, int nvec, hipStream_t s) const;
  // nvec complex columns interleaved: element (row, v) at row · ld + v (the
  // columns of a row contiguous, so one gathered cache line serves several
  // columns). Needs prepare_interleaved_action (compiled stream); returns
  // false (nothing launched) otherwise.
  bool apply_batch_interleaved(const cplx* x, cplx* y, size_t ld, int nvec, hipStream_t s) const;

  // Several evaluations of one real family at once (scan points solved in
  // lockstep). Supported when the family is real with Hermitian parts, its
  // sector needs no character phases, and the compiled compact stream and
  // the parts' stored diagonals exist (the stream is built on demand).
  bool scan_batch_supported(hipStream_t s) const;
  // per point b < nb: the gather-form amplitude table (scan_batch_amp_count
  // doubles at d_amps + b·count) and the per-row diagonal (re[N], im[N] at
  // d_diags + b·2N) of the family at coefs[b] (real coefficients)
  int scan_batch_amp_count() const;
  // the widest apply_scan_batch launch the lane tables fit (8 or 16)
  int scan_batch_max_width() const;
  void scan_batch_tables(const std::vector<std::vector<cplx>>& coefs, double* d_amps, double* d_diags,
                         hipStream_t s) const;
  // the two halves of scan_batch_tables for one point: its gather-form
  // amplitude table (host, scan_batch_amp_count entries) and its per-row
  // diagonal written to d_diag (device, re[N] then im[N])
  std::vector<double> scan_amp_table(const std::vector<cplx>& c) const;
  void scan_point_diagonal(const std::vector<cplx>& c, double* d_diag, hipStream_t s) const;
  // y_b = H(c_b) x_b for the lanes b < width (1, 2, 4, 8 or 16, at most
  // scan_batch_max_width) of an input
  // interleaved as x[row·width + b] (FP32 or FP64); y[b] == nullptr skips a
  // lane, and a nonzero *gates[b] (optional) voids it like the single-vector
  // gate. amps[b] / diags[b]: the lane's point tables from scan_batch_tables.
  void apply_scan_batch(const void* x, bool fp32, int width, const double* const* amps, const double* const* diags,
                        double* const* y, const int* const* gates, hipStream_t s) const;
  // Both precisions in one launch: lanes b < w32 read the FP32 interleaved
  // x32 (stride w32), lanes w32 <= b < w32 + w64 the FP64 interleaved x64
  // (stride w64); w32 in {0, 1, 2, 4, 8, 16}, w64 in {0, 1, 2, 4, 8}. Each
  // lane's result is the one apply_scan_batch gives it.
  void apply_scan_mixed(const float* x32, int w32, const double* x64, int w64, const double* const* amps,
                        const double* const* diags, double* const* y, const int* const* gates, hipStream_t s) const;

  // per-row Σ|H_ij| over j != i merged per column, plus |H_ii - shift|;
  // returns (trace, max row sum) — ‖A-μI‖₁ for Hermitian A
  std::pair<cplx, double> shifted_trace_and_one_norm(hipStream_t s) const;

  // number of entries the reference's assembly would store (merged per
  // (row, column), |v| > f64::EPSILON, diagonal included)
  size_t stored_entries(hipStream_t s) const;
  // lanes per row and H·v mode chosen by the autotuner
  int lanes_per_row() const;
  bool uses_image_table() const;
  bool uses_index_stream() const;

  // dense materialization (row-major, complex) on the device
  void materialize_dense(cplx* dense, hipStream_t s) const;

  // Storage format (OperatorBuilder::build(format), QMBED
  // src/operator/mod.rs:2091-2216): a stored format (csr, csc, dense, dia)
  // assembles the sector matrix on the device with the reference's stored
  // entries (merged per (row, column), |v| > f64::EPSILON) as CSR rows, and
  // the single-vector actions then run as sparse products over them;
  // matrix_free keeps the on-the-fly kernels. Dia answers unsupported
  // backend when the matrix is not diagonal-banded. A family re-assembles
  // the stored matrix after each evaluation.
  void use_format(MatrixFormat f, hipStream_t s);
  // the stored entries in row-major order (Operator::triplets,
  // src/operator/mod.rs:937-1003)
  struct Triplet {
    uint32_t row, col;
    cplx value;
  };
  std::vector<Triplet> triplets(hipStream_t s) const;
  // (re)assemble the stored matrix for the current evaluation
  void refresh_stored(hipStream_t s);

  // Operator families (QMBED src/operator/mod.rs:2710-2830, 3140-3230).
  bool parameterized() const { return !components_.empty(); }
  const std::vector<ComponentSpec>& components() const { return components_; }
  // named parameters → coefficients in component order (the reference's
  // resolve_coefficients: unknown names, missing required ones and
  // non-finite values are errors)
  std::vector<cplx> resolve_coefficients(const std::vector<std::pair<std::string, cplx>>& parameters) const;
  // Rebind the family to H_static + Σ_p c_p H_p: only the amplitude tables
  // change (basis, lookup and image tables are shared). Hermiticity of the
  // sum is known without a check when every part was Hermitian at build and
  // all c_p are real; otherwise the sector matrix is checked exactly.
  // Applies already queued on other streams must have completed.
  void evaluate_coefficients(const std::vector<cplx>& c, hipStream_t s);
  const std::vector<cplx>& coefficients() const { return coefficients_; }
  // the current sector matrix is Hermitian (the gather-form kernels apply)
  bool hermitian() const;

  friend class OperatorBuilder;
  struct DeviceTables;  // defined in hv_kernels.hip
  std::shared_ptr<DeviceTables> tables_;

 private:
  std::shared_ptr<const SpinBasis1D> basis_;
  std::vector<FlipGroup> diag_, off_;
  bool real_ = true;
  MatrixFormat format_ = MatrixFormat::MatrixFree;
  // families
  std::vector<ComponentSpec> components_;
  std::vector<cplx> coefficients_;
  std::vector<bool> part_hermitian_;  // static, then components
  bool parts_real_ = true;            // every part table real
  bool tuned_ = false;
};

// OperatorBuilder: QMBED src/operator/mod.rs:1941-2216.
class OperatorBuilder {
 public:
  static OperatorBuilder on(std::shared_ptr<const SpinBasis1D> basis) {
    OperatorBuilder b;
    b.basis_ = std::move(basis);
    return b;
  }
  OperatorBuilder& terms(const std::vector<OperatorSpec>& t) {
    terms_.insert(terms_.end(), t.begin(), t.end());
    return *this;
  }
  OperatorBuilder& term(const OperatorSpec& t) {
    terms_.push_back(t);
    return *this;
  }
  OperatorBuilder& checks(const AssemblyChecks& c) {
    checks_ = c;
    return *this;
  }
  // parameterized components (names nonempty and unique, defaults finite)
  OperatorBuilder& components(const std::vector<ComponentSpec>& c) {
    components_.insert(components_.end(), c.begin(), c.end());
    return *this;
  }
  std::shared_ptr<Operator> build(MatrixFormat format = MatrixFormat::MatrixFree) const;

 private:
  std::shared_ptr<const SpinBasis1D> basis_;
  std::vector<OperatorSpec> terms_;
  std::vector<ComponentSpec> components_;
  AssemblyChecks checks_;
};

// Validates symbols (QMBED src/operator/mod.rs:226-247 and the spin local
// action alphabet, src/basis/mod.rs:1160-1228).
void validate_terms(const std::vector<OperatorSpec>& terms, int sites);

// ------------------------------------------------------------------ solve --
enum class SpectrumTarget { SmallestAlgebraic, LargestAlgebraic, SmallestMagnitude, LargestMagnitude, BothEnds, Shift };

// EigshOptions: QMBED src/solve/mod.rs:349-449 (same defaults).
struct EigshOptions {
  int eigenpairs = 1;
  SpectrumTarget target = SpectrumTarget::SmallestAlgebraic;
  double shift = 0.0;
  std::optional<long> krylov_dimension;
  double tolerance = 1.0e-10;
  long max_iterations = 1000;
  uint64_t seed = 0;

  static EigshOptions smallest_algebraic(int k) {
    EigshOptions o;
    o.eigenpairs = k;
    return o;
  }
  EigshOptions& with_tolerance(double t) {
    tolerance = t;
    return *this;
  }
  EigshOptions& with_krylov_dimension(long d) {
    krylov_dimension = d;
    return *this;
  }
  EigshOptions& with_max_iterations(long m) {
    max_iterations = m;
    return *this;
  }
  EigshOptions& with_seed(uint64_t s) {
    seed = s;
    return *this;
  }
  void validate(size_t dimension) const;
};

// Precision of the matrix-vector product inside the Krylov loop. FP32 is a
// relaxed schedule: FP64 products until the wanted pairs' residual estimates
// fall below fp32_switch × tolerance, FP32 products after that (their errors
// enter the Ritz vectors through components of that size); a Ritz pair that
// then fails the FP64 test ends the window, which restarts in FP64 from its
// Ritz vectors. The Lanczos coefficients,
// reorthogonalization, tridiagonal solve and the final Rayleigh-quotient /
// residual test always stay FP64.
enum class MatvecPrecision { FP64 = 0, FP32 = 1 };

struct EigshControls {
  MatvecPrecision precision = MatvecPrecision::FP64;
  // check Ritz residual estimates every `check_every` steps inside a window
  int check_every = 2;
  // when true, stop inside the window as soon as the requested pairs have
  // converged (the reference always runs the full window)
  bool early_stop = true;
  // FP32 matvecs start once every wanted pair's residual estimate is below
  // fp32_switch × tolerance (relative to max(|θ|, 1))
  double fp32_switch = 1.0e4;
  // DGKS threshold η: a second reorthogonalization pass is taken when the
  // first leaves ‖w‖ <= η‖w₀‖ (the reference's constant is 1/√2; η = 1 takes
  // the second-pass path on nearly every step, which exercises it)
  double dgks_eta = 0.7071067811865476;
  // sector batches: at most this many concatenated (tile-padded) rows per
  // lockstep batch; 0 sizes the batches to the free device memory
  long batch_rows = 0;
  // steps launched between posting a read-back of the window's coefficients
  // and collecting it (the host then never waits for the device); 0 picks 1
  // for operators of at least 2^20 rows, whose steps outlast the host's work
  // on a read-back, and 2 below that
  int read_lag = 0;
  // scans from the snapshot start: points solved in lockstep per batch (one
  // shared H·v launch per step and 8 windows; 0 or 1 solves them one by one)
  int lockstep = 16;
  // lockstep scans: free slots take the next points once this many are free
  // (or none is running); 0 or lockstep and above: batches that start and
  // end together
  int lockstep_refill = 0;
  // warm-started lockstep scans: points are taken in ascending order of the
  // gap between the two lowest eigenvalues of the family projected on the
  // snapshot space (the hardest first), so that the points of a batch need
  // similar step counts; false keeps the given order. The results do not
  // depend on the order.
  bool lockstep_order = true;
  // return the eigenvectors (false: only eigenvalues and residuals, with no
  // device allocation per solve)
  bool keep_vectors = true;
};

struct Eigensystem {
  std::vector<double> eigenvalues;
  std::vector<std::vector<cplx>> eigenvectors;  // host copies (optional)
  std::vector<double> residuals;
  long iterations = 0;
  long reorthogonalization_passes = 0;
  long conditional_second_passes = 0;
  bool converged = true;
  long fp32_iterations = 0;  // matvecs spent in the FP32 stage
  long fp64_iterations = 0;  // matvecs spent in FP64 (incl. polish)
};

// Lockstep Lanczos windows for several evaluations of one real family (scan
// points): one window per point, all stepping together, with
// apply_scan_batch launches for every window's H·v (the windows' next
// vectors interleaved, so the gathers of all points share cache lines).
// Each window runs the single-window algorithm (same checks, FP32 schedule,
// DGKS passes) at the default window size. Each step packs the next vectors
// of the windows still stepping into consecutive lanes, grouped by each
// window's product precision (launch width 1, 2, 4, 8 or 16), so converged
// windows cost nothing. lanes64 / lanes32: the interleaved inputs (grown as
// needed, kept by the caller between calls). A point whose window fails
// comes back with ok == false, for the caller to solve on its own.
struct LockstepPoint {
  const double* start;  // FP64 start vector (device), normalized here
  const double* amps;   // the point's amplitude table (device)
  const double* diag;   // the point's per-row diagonal (device)
  double* vector = nullptr;  // optional: receives the accepted Ritz vector (device, N)
};
struct LockstepResult {
  Eigensystem es;
  bool ok = false;
};
std::vector<LockstepResult> eigsh_lockstep(const Operator& op, const EigshOptions& opts, const EigshControls& ctl,
                                           const std::vector<LockstepPoint>& pts, DeviceArray<double>& lanes64,
                                           DeviceArray<float>& lanes32, hipStream_t s);
// Continuous lockstep: `slots` windows (1 to 16) step together, and a slot
// whose window has finished is refilled at once with the next point, so the
// launches stay wide until the points run out. feed.next(slot, pt) fills pt
// with the next point's device tables and start (written in stream order
// into the slot's own areas, which the slot's previous point no longer
// reads) and returns its index, or -1 when none is left; feed.done(index,
// result) hands over each point's result (ok == false: the caller solves
// that point on its own). Every window runs the single-window algorithm, so
// each point's result is that of its one-by-one solve.
struct LockstepFeed {
  std::function<int(int slot, LockstepPoint& pt)> next;
  std::function<void(int index, LockstepResult&& result)> done;
  // optional: called before the next() calls of one refill with the number
  // of free slots, so the caller can prepare that many points at once
  std::function<void(int free)> prepare;
  // free slots are refilled once at least `group` of them are free, or when
  // no window is left running (group >= slots: batches that start and end
  // together)
  int group = 1;
};
void eigsh_lockstep_stream(const Operator& op, const EigshOptions& opts, const EigshControls& ctl, int slots,
                           const LockstepFeed& feed, DeviceArray<double>& lanes64, DeviceArray<float>& lanes32,
                           hipStream_t s);

// eigsh: QMBED src/solve/mod.rs:2335-2452 (dispatch), 1116-1456 (Lanczos).
// `initial` is a host vector of length dim (nullptr: deterministic start).
Eigensystem eigsh(const Operator& op, const EigshOptions& opts, const cplx* initial = nullptr,
                  const EigshControls& ctl = EigshControls(), bool want_vectors = false, hipStream_t s = nullptr);

// eigh: the complete dense spectrum of a sector (QMBED src/solve/mod.rs:586-620)
Eigensystem eigh_all(const Operator& op, bool want_vectors, hipStream_t s);

// eigsh_with_initial_subspace: QMBED src/solve/mod.rs:2385-2410. Every
// vector is validated (length, finite entries, nonzero norm); no vector runs
// eigsh, one runs eigsh from it, several run the thick restart with the whole
// subspace kept (small problems take the dense route either way).
Eigensystem eigsh_subspace(const Operator& op, const EigshOptions& opts, const std::vector<std::vector<cplx>>& sub,
                           const EigshControls& ctl, bool want_vectors, hipStream_t s);

// EigshWorkspace: QMBED src/solve/mod.rs:477-526, the converged invariant
// subspace of the previous solve of a related operator (host vectors).
struct EigshWorkspace {
  size_t dimension = 0;
  std::vector<std::vector<cplx>> subspace;
  void clear() {
    dimension = 0;
    subspace.clear();
  }
  // set_initial_subspace: every vector validated against `dim`
  void set(size_t dim, std::vector<std::vector<cplx>> vectors);
};
// eigsh_with_workspace: QMBED src/solve/mod.rs:2419-2444. A workspace of the
// operator's dimension with more than one vector starts the large-window
// route (no krylov_dimension, or at least 96) from balanced_subspace_start
// (:708-727) and every other case from the whole subspace; the workspace
// then holds this solve's eigenvectors.
Eigensystem eigsh_workspace(const Operator& op, const EigshOptions& opts, EigshWorkspace& ws,
                            const EigshControls& ctl, hipStream_t s);
// balanced_subspace_start: Σ_j e^{2πij/J} v_j / √J normalized (v_0 when the
// sum vanishes)
std::vector<cplx> balanced_subspace_start(const std::vector<std::vector<cplx>>& sub);
// validate_eigsh_initial: QMBED src/solve/mod.rs:671-692
void validate_eigsh_initial(const std::vector<cplx>& v, size_t dimension);

// Many sectors at once (sector_batch.hip): the Krylov-route Hermitian
// sectors run one fully reorthogonalized Lanczos window each, in lockstep, so
// every launch of a step covers all of them; sectors that do not converge
// within the common window, dense-route and non-Hermitian sectors are solved
// one by one afterwards. Results are in input order; each sector's pairs pass
// the same FP64 residual test as eigsh. FP64 matvecs only.
struct SectorBatchStats {
  int sectors = 0;      // sectors requested
  int batched = 0;      // converged inside the lockstep batch
  int fallback = 0;     // batched sectors finished by the single-sector solver
  long steps = 0;       // lockstep steps launched
  int ritz_rounds = 0;  // Ritz tests (each one synchronization)
  int window = 0;       // common Krylov window
  int batches = 0;      // consecutive lockstep batches (device memory)
  double seconds = 0;   // wall time of the call
};
std::vector<Eigensystem> eigsh_sectors(const std::vector<Operator*>& ops, const EigshOptions& opts,
                                       const EigshControls& ctl, bool want_vectors, hipStream_t s,
                                       SectorBatchStats* stats = nullptr);

// Device-side variant: initial vector already on the device (real, FP64);
// returned eigenvectors (if requested) are written to `vectors_out` (device,
// real FP64, eigenpairs × dim) for real operators.
Eigensystem eigsh_device(const Operator& op, const EigshOptions& opts, const double* d_initial,
                         const EigshControls& ctl, double* d_vectors_out, hipStream_t s);

// Device start vector (interleaved complex when initial_complex or the
// operator is complex; nullptr: deterministic start); the eigenvectors stay
// on the device in `vectors` (eigenpairs columns, complex when
// complex_vectors).
Eigensystem eigsh_on_device(const Operator& op, const EigshOptions& opts, const double* d_initial,
                            bool initial_complex, const EigshControls& ctl, DeviceArray<double>& vectors,
                            bool& complex_vectors, hipStream_t s);

// Krylov time evolution: QMBED src/solve/mod.rs:2454-3454.
struct EvolutionOptions {
  std::vector<double> times;
  long krylov_dimension = 30;
  double tolerance = 1.0e-10;
  long max_substeps = 10000;
  bool hamiltonian = true;
  void validate() const;
};

struct EvolutionDiagnostics {
  long lanczos_projections = 0;
  long matrix_vector_products = 0;
  long real_lanczos_projections = 0;
  long real_matrix_vector_products = 0;
  long accepted_substeps = 0;
  long rejected_trial_intervals = 0;
  double maximum_estimated_error = 0.0;
};

// Builds (once) the compiled index stream an operator needs for
// apply_batch_interleaved; false when the sector does not fit it.
bool prepare_interleaved_action(const Operator& op, hipStream_t s);

// Device-resident evolution: d_cols holds ncol complex columns of length
// dim ([column][row]); d_out receives [time][column][row]. One host
// synchronization per batched Lanczos projection.
void evolve_batch_device(const Operator& op, const cplx* d_cols, int ncol, const EvolutionOptions& opts,
                         EvolutionDiagnostics* diag, cplx* d_out, hipStream_t s);

// Evolve `ncol` host columns (column-major, dim each); output is
// [time][column][dim] on the host.
std::vector<std::vector<std::vector<cplx>>> evolve_batch(const Operator& op, const std::vector<std::vector<cplx>>& columns,
                                                         const EvolutionOptions& opts, EvolutionDiagnostics* diag,
                                                         hipStream_t s);

// Truncated-Taylor exponential-action plan: QMBED src/solve/mod.rs:116-248.
class ExpmActionPlan {
 public:
  ExpmActionPlan(std::shared_ptr<const Operator> op, cplx coefficient, int max_degree, double tolerance,
                 long max_substeps, hipStream_t s);
  size_t dimension() const { return op_->dim(); }
  int degree() const { return degree_; }
  long scaling() const { return scaling_; }
  cplx shift() const { return shift_; }
  double one_norm() const { return one_norm_; }
  cplx coefficient() const { return coefficient_; }
  // in-place on device columns (nvec × dim, complex)
  void apply_device(cplx* d_vectors, int nvec, hipStream_t s) const;
  std::vector<std::vector<cplx>> apply(const std::vector<std::vector<cplx>>& vectors, hipStream_t s) const;

 private:
  std::shared_ptr<const Operator> op_;
  cplx coefficient_, shift_;
  double one_norm_ = 0;
  int degree_ = 0;
  long scaling_ = 1;
  double tolerance_ = 0;
};

// Host symmetric tridiagonal eigensolver (implicit QL with Wilkinson shifts);
// eigenvalues ascending; `vecs` (row-major m×m, columns = eigenvectors) only
// when want_vectors.
void tridiagonal_eigen(const std::vector<double>& alpha, const std::vector<double>& beta, std::vector<double>& values,
                       std::vector<double>* vecs);

// Host dense Hermitian eigen-decomposition (Householder tridiagonalization,
// implicit QL, back-transformation) of a row-major n×n matrix: values
// ascending, vectors row-major with column j the unit eigenvector of
// values[j].
void hermitian_eigen(int n, std::vector<cplx> a, std::vector<double>& values, std::vector<cplx>& vectors);

// Index selection per spectral target (QMBED src/solve/mod.rs:1002-1036).
std::vector<int> select_indices(const std::vector<double>& values, SpectrumTarget target, int count, double shift = 0.0);
// eigenvalues lo..hi-1 (ascending, 0 = smallest) of a symmetric tridiagonal
// by bisection
std::vector<double> tridiagonal_eigenvalue_range(const std::vector<double>& alpha, const std::vector<double>& beta,
                                                 int lo, int hi);
// The Ritz values a target selects from the tridiagonal (alpha, beta), in
// select_indices order; with `vectors`, their unit eigenvectors too.
// Extreme targets (smallest/largest algebraic, both ends, largest magnitude)
// find only the k lowest and/or highest values by bisection and their
// vectors by inverse iteration (the full QL decomposition when selected
// values nearly coincide); the other targets take the full QL spectrum.
std::vector<double> selected_ritz_values(const std::vector<double>& alpha, const std::vector<double>& beta,
                                         SpectrumTarget target, int k, double shift,
                                         std::vector<std::vector<double>>* vectors = nullptr);

// Deterministic start (QMBED src/solve/mod.rs:729-742), written on the device.
// complex: interleaved (value, 0) pairs
void deterministic_start_device(double* d_out, size_t n, uint64_t seed, hipStream_t s, bool complex = false);

int device_count();

}  // namespace qmbed
