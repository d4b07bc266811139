/*
 * qmbed_hip — C ABI of the MI355X-native spin-sector H·v / eigsh / Krylov
 * exponential engine.
 *
 * The first five entry points are drop-ins for the reference's C ABI
 * (QMBED bindings/capi/src/lib.rs), with the same JSON request and response
 * schemas for the spin-basis eigsh / evolve / exponential-action commands and
 * the same ownership rule (responses are heap strings released exactly once
 * with qmbed_string_free). The qmbed_hip_* functions are a typed fast path
 * for device-resident vectors; they have no counterpart in the reference.
 */
#ifndef QMBED_HIP_H
#define QMBED_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference-compatible JSON protocol ------------------------------ */

/* Replaces qmbed_run_json (bindings/capi/src/lib.rs:6089-6118): one solve
 * request {"basis","terms","format","solver"} -> {"status":"ok","result":{...}}
 * or {"status":"error","error":"..."}. */
char* qmbed_run_json(const char* request);

/* Replaces qmbed_command_json (bindings/capi/src/lib.rs:6120-6151): tagged
 * "operation" commands describe_basis, create_model, eigsh_model,
 * evolve_model, create_expm_action, apply_expm_action, release_model,
 * release_expm_action. Other operations answer an error. */
char* qmbed_command_json(const char* request);

/* ABI-stable complex scalar of the coefficient callbacks
 * (QmbedComplex64, bindings/capi/src/lib.rs:452-458). */
typedef struct QmbedComplex64 {
  double real;
  double imaginary;
} QmbedComplex64;

/* QmbedDriveCallback (bindings/capi/src/lib.rs:460-469): fill `count`
 * ordered component coefficients for one physical time; nonzero aborts. */
typedef int32_t (*QmbedDriveCallback)(void* context, double time, QmbedComplex64* coefficients, size_t count);

/* Replaces qmbed_evolve_model_with_drive_json
 * (bindings/capi/src/lib.rs:6153-6179); exported so existing loaders resolve
 * it. Driven evolution is outside this engine: a null request answers
 * "request pointer is null", a null callback "drive callback is null",
 * anything else an unsupported-backend error. */
char* qmbed_evolve_model_with_drive_json(const char* request, QmbedDriveCallback callback, void* context);

/* QuSpin-compatible user-basis callback layouts
 * (bindings/capi/src/lib.rs:471-521). */
typedef struct QmbedUserOpResult32 {
  QmbedComplex64 matrix_element;
  uint32_t state;
} QmbedUserOpResult32;
typedef struct QmbedUserOpResult64 {
  QmbedComplex64 matrix_element;
  uint64_t state;
} QmbedUserOpResult64;
typedef int32_t (*QmbedUserOp32Callback)(QmbedUserOpResult32* result, int8_t op, int32_t site, int32_t sites,
                                         const uint32_t* arguments);
typedef uint32_t (*QmbedUserNextState32Callback)(uint32_t state, uint32_t counter, uint32_t sites,
                                                 const uint32_t* arguments);
typedef uint32_t (*QmbedUserPreCheck32Callback)(uint32_t state, uint32_t sites, const uint32_t* arguments);
typedef uint32_t (*QmbedUserMap32Callback)(uint32_t state, int32_t sites, int8_t* sign, const uint32_t* arguments);
typedef int32_t (*QmbedUserOp64Callback)(QmbedUserOpResult64* result, int8_t op, int32_t site, int32_t sites,
                                         const uint64_t* arguments);
typedef uint64_t (*QmbedUserNextState64Callback)(uint64_t state, uint64_t counter, uint64_t sites,
                                                 const uint64_t* arguments);
typedef uint64_t (*QmbedUserPreCheck64Callback)(uint64_t state, uint64_t sites, const uint64_t* arguments);
typedef uint64_t (*QmbedUserMap64Callback)(uint64_t state, int32_t sites, int8_t* sign, const uint64_t* arguments);

/* Replace qmbed_register_user_basis_{32,64}_json
 * (bindings/capi/src/lib.rs:6181-6288), with the reference's six
 * parameters; exported for loader compatibility. User bases are outside
 * this engine: after the reference's own argument checks (null request,
 * null operator callback, null map array with map_count > 0) they answer an
 * unsupported-backend error. */
char* qmbed_register_user_basis_32_json(const char* request, QmbedUserOp32Callback op,
                                        QmbedUserNextState32Callback next_state,
                                        QmbedUserPreCheck32Callback pre_check,
                                        const QmbedUserMap32Callback* map_callbacks, size_t map_count);
char* qmbed_register_user_basis_64_json(const char* request, QmbedUserOp64Callback op,
                                        QmbedUserNextState64Callback next_state,
                                        QmbedUserPreCheck64Callback pre_check,
                                        const QmbedUserMap64Callback* map_callbacks, size_t map_count);

/* Replaces qmbed_string_free (bindings/capi/src/lib.rs:6290-6297). NULL is
 * allowed. */
void qmbed_string_free(char* response);

/* ---- typed fast path (no counterpart in the reference) ---------------- */

typedef struct qmbed_hip_model qmbed_hip_model;

enum { QMBED_F64 = 0, QMBED_F32 = 1, QMBED_C128 = 2, QMBED_C64 = 3 };
/* Mixed-precision actions (compiled-stream mode only; other modes answer
 * -3): FP32 input, FP64 accumulation and output, real (x float, y double)
 * or complex (x complex64, y complex128). The FP32 Lanczos schedule's
 * product. */
enum { QMBED_F32_TO_F64 = 4, QMBED_C64_TO_C128 = 5 };

/* model_json: {"basis":{...},"terms":[...],"components":[...]?,
 * "checks":{...}?,"format":"csr"|"csc"|"dense"|"dia"|"matrix_free"?} as in
 * create_model (components: [{"name", "terms", "default": [re, im]?}],
 * bindings/capi/src/lib.rs:783-787). Builds the sector on the GPU and
 * compiles the operator; with a stored format the products run on the
 * stored matrix assembled on the GPU.
 * On failure returns NULL and stores an error string (free with
 * qmbed_string_free) in *error when error is non-NULL. */
qmbed_hip_model* qmbed_hip_model_create(const char* model_json, char** error);
void qmbed_hip_model_free(qmbed_hip_model* model);
long qmbed_hip_model_dimension(const qmbed_hip_model* model);
int qmbed_hip_model_is_real(const qmbed_hip_model* model);
/* host copies of the sorted representatives and their orbit sizes */
int qmbed_hip_model_states(const qmbed_hip_model* model, uint32_t* states, uint8_t* orbits);
// the representatives as 64-bit words in row order (sectors of any width;
// qmbed_hip_model_states answers -2 for the wide sectors of L >= 32)
int qmbed_hip_model_states64(const qmbed_hip_model* model, uint64_t* states, uint8_t* orbits);
/* entries the reference's assembly would store for this operator (merged
 * per (row, column), |v| > machine epsilon, diagonal included); -1 on error */
long qmbed_hip_model_nnz(const qmbed_hip_model* model);
/* H·v configuration picked by the autotuner: lanes per row and the mode in
 * *image_table: 0 orbit walk + sorted-table search, 1 image table,
 * 2 compiled index stream */
int qmbed_hip_model_hv_config(const qmbed_hip_model* model, int* lanes_per_row, int* image_table);
/* the model's HIP stream (hipStream_t) */
void* qmbed_hip_model_stream(const qmbed_hip_model* model);
/* y = H x on device pointers (dtype QMBED_*); stream NULL = model stream */
int qmbed_hip_apply_device(const qmbed_hip_model* model, int dtype, const void* x, void* y, void* stream);
/* y = H x on host arrays (copies through device memory) */
int qmbed_hip_apply_host(const qmbed_hip_model* model, int dtype, const void* x, void* y);
/* average milliseconds of `iters` back-to-back H·v launches (HIP events on
 * the model stream) */
double qmbed_hip_time_apply(const qmbed_hip_model* model, int dtype, const void* x, void* y, int iters);
/* Average milliseconds of `iters` launches of the batched scan H·v (the
 * kernel of lockstep scan solves) over `width` lanes (1, 2, 4, 8 or 16)
 * of FP32 (fp32 != 0) or FP64 interleaved inputs, each lane its own point
 * of the family (points: width rows of one real coefficient per
 * component), timed with HIP events on the model stream after two
 * untimed launches. -2 when the model does not support batched scan
 * launches of that width, -1 on failure. No counterpart in the reference. */
double qmbed_hip_time_scan_batch(qmbed_hip_model* model, int width, int fp32, const double* points, int iters);
/* y_b = H(c_b) x_b through the same batched scan H·v for `width` lanes (1, 2,
 * 4, 8 or 16): points as above, x and y `width` host vectors of the
 * dimension (column b at offset b·n), the inputs rounded to FP32 when
 * fp32 != 0. 0 on success, -2 when the model does not support batched scan
 * launches of that width, -1 on failure. No counterpart in the reference. */
int qmbed_hip_apply_scan_batch(qmbed_hip_model* model, int width, int fp32, const double* points, const double* x,
                               double* y);
/* Launch timer (no counterpart in the reference): on != 0 clears the
 * records and brackets every later batched scan H·v launch with two HIP
 * events on its stream; 0 stops recording. The report answers
 * {"status":"ok","result":{"<kernel instance>":{"launches","ms_total",
 * "ms_mean"},...}} over the launches recorded since the last switch-on
 * (it waits for them). */
int qmbed_hip_launch_timer(int on);
char* qmbed_hip_launch_timer_report(void);

/* eigsh with a SolverRequest JSON; response as qmbed_run_json's result plus
 * fp32_iterations / fp64_iterations / reorthogonalization counters. Accepts
 * the extra solver keys "matvec_precision" ("fp64"|"fp32") and
 * "early_stop" (bool), "check_every" (steps between Ritz-estimate checks,
 * 1-64, default 2), "read_lag" (steps launched between posting a read-back
 * and collecting it, 0-8; 0, the default, picks 1 from 2^20 rows on and 2
 * below) and "reorthogonalization_eta" (DGKS threshold, (0, 1]). */
char* qmbed_hip_eigsh_json(const qmbed_hip_model* model, const char* solver_json);

/* The solver JSON of qmbed_hip_eigsh_json also accepts "initial_subspace":
 * [[[re, im], ...], ...] (several start vectors; not with initial_vector),
 * solved as eigsh_with_initial_subspace (src/solve/mod.rs:2385-2410): one
 * vector takes the single-start route, several the thick restart, which
 * keeps them as a subspace. Explicit krylov_dimension below 96 takes the
 * thick restart too (src/solve/mod.rs:1644-2196, 2299-2327). */

/* EigshWorkspace (src/solve/mod.rs:477-526): the converged eigenvectors of
 * the previous solve of a related operator, host-resident. */
typedef struct qmbed_hip_workspace qmbed_hip_workspace;
qmbed_hip_workspace* qmbed_hip_workspace_create(void);
void qmbed_hip_workspace_free(qmbed_hip_workspace* workspace);
long qmbed_hip_workspace_dimension(const qmbed_hip_workspace* workspace);
int qmbed_hip_workspace_vectors(const qmbed_hip_workspace* workspace);
/* set_initial_subspace: `count` complex vectors of `dimension` entries
 * (interleaved re, im), each validated as an eigsh initial vector. */
char* qmbed_hip_workspace_set(qmbed_hip_workspace* workspace, long dimension, int count, const double* vectors);
/* eigsh_with_workspace (src/solve/mod.rs:2419-2444): with a workspace of
 * the model's dimension holding more than one vector, the large-window
 * route (no krylov_dimension, or one of at least 96) starts from
 * balanced_subspace_start (:708-727), every other case from the whole
 * subspace; the workspace then holds this solve's eigenvectors. Response as
 * qmbed_hip_eigsh_json. */
char* qmbed_hip_eigsh_workspace_json(const qmbed_hip_model* model, qmbed_hip_workspace* workspace,
                                     const char* solver_json);

/* Solve `count` models (typically the symmetry sectors of one Hamiltonian)
 * with one SolverRequest in a lockstep batch: every Lanczos launch covers all
 * unconverged sectors. Response: {"kind":"eigensystems","results":[one
 * qmbed_hip_eigsh_json result per model, in order],"batch":{"sectors",
 * "batched","fallback","steps","ritz_rounds","window","seconds"}}. Every
 * sector starts from the deterministic start vector (no initial_vector);
 * FP64 matvecs. No counterpart in the reference's ABI: its callers loop over
 * eigsh per sector (the same results, within the solver tolerance); the
 * JSON command {"operation":"eigsh_models","handles":[...],"solver":{...}}
 * is the same call on registered handles. Batches are sized to the free
 * device memory (the solver key "batch_rows" caps the concatenated rows of
 * one batch; "batches" in the answer counts them). */
char* qmbed_hip_eigsh_batch_json(qmbed_hip_model* const* models, int count, const char* solver_json);

/* Krylov evolution of device-resident states: `states` holds ncols complex128
 * columns of length dim ([column][row]); `out` receives
 * [time][column][row] (times × ncols × dim complex128). evolution_json:
 * {"times":[...], "krylov_dimension" (default 30, EvolutionOptions::new),
 * "tolerance" (1e-10), "max_substeps" (10000), "parameters" (families)}.
 * Answers {"kind":"evolution","dimension","times"} plus the reference's
 * EvolutionDiagnostics counters. Replaces the vector transfer of
 * evolve_model (QMBED capi evolve_model_json; evolve_batch,
 * src/solve/mod.rs:4141-4170) for callers whose states live on the GPU. */
char* qmbed_hip_evolve_device(qmbed_hip_model* model, const void* states, int ncols, void* out,
                              const char* evolution_json);

/* Evaluate a parameterized model (created with "components") at
 * {"name": [re, im], ...}: H = H_static + sum_p c_p H_p, defaults filling
 * absent names. Later applies and solves use these values. Answers the
 * coefficients in component order, or the reference's error text (unknown
 * or missing parameter, non-finite value, non-Hermitian evaluation). No
 * counterpart in the reference's ABI (its commands take "parameters" per
 * call, which eigsh_json also accepts). */
char* qmbed_hip_model_set_parameters(qmbed_hip_model* model, const char* parameters_json);

/* ---- Learned Lanczos warm start for parameter scans ----
 * No counterpart in the reference (its scan reuse is EigshWorkspace's
 * balanced subspace start, src/solve/mod.rs:477-526 and :2419-2444). A scan
 * is bound to a model created with "components" (one real coefficient per
 * component is one scan coordinate).
 * options_json: {"snapshots":32,"hidden":[64,64],"epochs":4000,
 * "learning_rate":0.002,"final_learning_rate":<learning_rate>,"seed":0,
 * "targets":"projections","samples":512,"provisional":true,"space":"greedy"},
 * any subset, NULL for all
 * defaults (the learning rate decays exponentially to final_learning_rate;
 * equal values keep it constant). targets "projections": the network learns
 * the coefficients of the training ground states over the snapshots;
 * "continuation": it learns the lowest eigenvector of the family projected
 * on the snapshots at the training points and `samples` seeded points of
 * their bounding box; "correction": it learns the step from that lowest
 * eigenvector y(c) to the training ground state's normalized projection on
 * the snapshots, and its start is S·(y(c) + MLP(c)). provisional: training points after the first
 * `snapshots` start from the projection on the snapshot space of those
 * already solved. space "greedy": `snapshots` training ground states chosen
 * by pivoted Cholesky of their Gram matrix; "principal": the `snapshots`
 * leading principal components of all training ground states. */
typedef struct qmbed_hip_scan qmbed_hip_scan;
qmbed_hip_scan* qmbed_hip_scan_create(qmbed_hip_model* model, const char* options_json, char** error);
void qmbed_hip_scan_free(qmbed_hip_scan* scan);
/* Solve the n training points (row-major n × components) with the plain
 * start and solver_json (as qmbed_hip_eigsh_json; eigenpairs is taken as 1;
 * in lockstep batches of up to "lockstep" points as qmbed_hip_scan_solve,
 * results equal to one-by-one solves), select the snapshot ground states and
 * train the FP32 network on the GPU.
 * Answers a JSON report (energies, iterations, snapshot indices, timings,
 * first/last training loss). */
char* qmbed_hip_scan_train(qmbed_hip_scan* scan, const double* points, int n, const char* solver_json);
/* Ground states of n points. warm = 0: the plain deterministic start;
 * 1: the network's start S·MLP(c) (S·(y(c) + MLP(c)) under correction
 * targets); 2: the snapshot Rayleigh-Ritz start
 * S·y, y the lowest eigenvector of Sᵀ H(c) S (eigenvector continuation over
 * the learned snapshots; the projected parts are formed once at training).
 * Answers energies, residuals, iterations (total, FP32-matvec and FP64),
 * per-point seconds and fallback count. The points are solved (any start)
 * in lockstep batches of up to "lockstep" (solver key, 0-16, default 16; 0
 * or 1: one by one): one Lanczos window per point; per step, shared H·v
 * launches over the interleaved vectors of the windows still stepping (each
 * point with its own couplings and each window at its own product
 * precision: up to 16 FP32 and 8 FP64 vectors in one launch), and every
 * other phase of the step in one launch for all of them. Each window runs
 * the single-window algorithm, so results equal one-by-one solves bit for
 * bit; a point whose window fails is solved on its own. By default the
 * points run in batches that start and end together; the solver key
 * "refill_group" (0-16) lets free slots take the next points once that many
 * are free (the starts of one refill formed together), and "refill" (bool)
 * is refill_group 1: a finished window's slot goes to the next point at
 * once. Needs a real family with Hermitian parts, no momentum
 * phases, the compact compiled stream, more than 256 rows and no explicit
 * window below 96; otherwise the points are solved one by one. */
char* qmbed_hip_scan_solve(qmbed_hip_scan* scan, const double* points, int n, const char* solver_json, int warm);
/* Network outputs (n × snapshots, row-major) of n points, in coefficient
 * units (under correction targets: the learned correction, without y(c));
 * returns the snapshot count, or -1 on failure. */
int qmbed_hip_scan_predict(qmbed_hip_scan* scan, const double* points, int n, double* coefficients);

/* The FP32 network kernels themselves (host buffers; for checks against a
 * CPU implementation). Shape in → h1 (tanh) → h2 (tanh) → out, parameters
 * W1[h1][in], b1, W2[h2][h1], b2, W3[out][h2], b3; loss 1/(2n) Σ‖y − t‖².
 * mlp_train with epochs == 0 writes one gradient into grad and the loss
 * into loss_history[0]; with epochs > 0 it runs full-batch Adam, updates
 * params in place and writes every epoch's loss into loss_history
 * (optional). Return 0 on success, -1 on failure. */
int qmbed_hip_mlp_parameter_count(int in, int h1, int h2, int out);
int qmbed_hip_mlp_forward(int in, int h1, int h2, int out, const float* params, const float* x, int n, float* y);
int qmbed_hip_mlp_train(int in, int h1, int h2, int out, float* params, const float* x, const float* t, int n,
                        int epochs, float learning_rate, float* grad, float* loss_history);

/* C = Aᵀ B for column-major device arrays A (n × p) and B (n × q) of real
 * doubles, n ≫ p, q: the tall-skinny product of the snapshot projections
 * (no counterpart in the reference; exported for tests and device callers).
 * C is p × q, column-major, on the device; the call synchronizes. Return 0
 * on success, -1 on failure. */
int qmbed_hip_gemm_atb(const double* A, size_t n, int p, const double* B, int q, double* C);
/* qmbed_hip_gemm_atb without device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * per-block partial products live in a static device array of 4 MiB
 * (524,288 doubles) per device, held until the process ends: every p, q <= 64
 * fits, and larger shapes up to that size; beyond it the call answers -2
 * and launches nothing. Same kernels and result as qmbed_hip_gemm_atb; the
 * call synchronizes the device. Calls that overlap on one device race on
 * the array. */
int qmbed_hip_gemm_atb_noalloc(const double* A, size_t n, int p, const double* B, int q, double* C);

void* qmbed_hip_malloc(size_t bytes);
void qmbed_hip_free(void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int qmbed_hip_memcpy(void* dst, const void* src, size_t bytes, int kind);
int qmbed_hip_synchronize(void);
int qmbed_hip_device_count(void);

/* Basis manifests: the versioned zip archive of ordered basis states
 * (src/archive.rs:333-362, save_basis_zip over BasisArchive::new and
 * insert_metadata, :246-331). Host-only; no device work.
 * request: {"path": str, "width_bits": w, "states": [decimal strings or
 * integers], "metadata": {key: str}}. When `states` is non-null its n 64-bit
 * words are the states (and the request carries no "states"). The states are
 * kept in the given row order; they must be unique and fit w bits. Answers
 * {"status":"ok","result":{"path","width_bits","count"}} or the reference's
 * error text ("archive error: …", "unsupported backend: integer needs …").
 * Free the answer with qmbed_string_free. */
char* qmbed_hip_basis_archive_save(const char* request, const uint64_t* states, size_t n);
/* Read one (src/archive.rs:364-416, load_basis_zip): answers
 * {"width_bits","count","fits_u64","written","metadata":{…}} plus
 * "states": [decimal strings] when `decimal` is nonzero. The states are
 * written to `states` (row order) when it is non-null, every state fits 64
 * bits and count <= capacity ("written": true); call with capacity 0 first to
 * learn the count. Stored and deflated entries are read, zip64 fields too. */
char* qmbed_hip_basis_archive_load(const char* path, uint64_t* states, size_t capacity, int decimal);

#ifdef __cplusplus
}
#endif

#endif /* QMBED_HIP_H */
