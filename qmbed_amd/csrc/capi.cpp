// C ABI: the JSON protocol of qmbed-capi (QMBED bindings/capi/src/lib.rs:
// qmbed_run_json :6089-6118, qmbed_command_json :6120-6151,
// qmbed_string_free :6290-6297) for the spin-sector eigsh / evolve /
// exponential-action path, plus a typed device-pointer fast path.
//
// Every entry point catches all exceptions and answers
// {"status":"error","error":"..."}; responses are malloc'd C strings owned by
// the caller until qmbed_string_free.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <optional>
#include <set>
#include <sstream>

#include "../../include/qmbed_hip.h"
#include "archive.hpp"
#include "internal.hpp"
#include "lanczos_kernels.hpp"
#include "json.hpp"
#include "scan.hpp"

namespace qmbed {

namespace {

using json::Value;

char* dup_c(const std::string& s) {
  char* p = static_cast<char*>(std::malloc(s.size() + 1));
  std::memcpy(p, s.c_str(), s.size() + 1);
  return p;
}

std::string error_json(const std::string& msg) {
  std::string o = "{\"status\":\"error\",\"error\":";
  json::write_string(o, msg);
  o += "}";
  return o;
}

// ---------------------------------------------------------- request parse --
[[noreturn]] void bad(const std::string& what) { throw invalid_options(what); }

const Value& need(const Value& obj, const char* key, const std::string& ctx) {
  const Value* v = obj.get(key);
  if (!v) bad(ctx + ": missing field `" + key + "`");
  return *v;
}

// integers within ±2^53 (exact in a double, and far inside long's range)
long as_long(const Value& v, const std::string& what) {
  if (v.kind != Value::Number || !std::isfinite(v.num) || std::floor(v.num) != v.num)
    bad(what + ": expected an integer");
  if (std::fabs(v.num) > 9007199254740992.0) bad(what + ": integer out of range");
  return long(v.num);
}

int as_int(const Value& v, const std::string& what) {
  const long x = as_long(v, what);
  if (x < long(INT32_MIN) || x > long(INT32_MAX)) bad(what + ": integer out of range");
  return int(x);
}

uint64_t as_u64(const Value& v, const std::string& what) {
  if (v.kind != Value::Number) bad(what + ": expected an integer");
  if (v.raw.empty() || v.raw.find_first_not_of("0123456789") != std::string::npos)
    bad(what + ": expected a non-negative integer");
  errno = 0;
  const unsigned long long u = std::strtoull(v.raw.c_str(), nullptr, 10);
  if (errno == ERANGE) bad(what + ": integer out of range");
  return uint64_t(u);
}

double as_double(const Value& v, const std::string& what) {
  if (v.kind != Value::Number) bad(what + ": expected a number");
  return v.num;
}

bool as_bool(const Value* v, bool dflt) {
  if (!v || v->is_null()) return dflt;
  if (v->kind != Value::Bool) bad("expected a boolean");
  return v->b;
}

cplx as_complex(const Value& v, const std::string& what) {
  if (v.kind != Value::Arr || v.arr->size() != 2) bad(what + ": expected [re, im]");
  return cplx(as_double((*v.arr)[0], what), as_double((*v.arr)[1], what));
}

std::vector<cplx> as_cvec(const Value& v, const std::string& what) {
  if (v.kind != Value::Arr) bad(what + ": expected an array of [re, im]");
  std::vector<cplx> out;
  out.reserve(v.arr->size());
  for (auto& e : *v.arr) out.push_back(as_complex(e, what));
  return out;
}

// BasisRequest::Spin (capi :620-640, :719-721; builder :2849-2920)
SpinBasisSpec parse_basis(const Value& b) {
  if (b.kind != Value::Obj) bad("invalid binding request: basis must be an object");
  const Value* kind = b.get("kind");
  if (!kind || kind->kind != Value::String) bad("invalid binding request: missing field `kind`");
  if (kind->raw != "spin")
    throw unsupported_backend("the MI355X engine implements spin bases; basis kind `" + kind->raw + "` is not built");
  SpinBasisSpec s;
  s.sites = as_int(need(b, "sites", "basis"), "sites");
  if (s.sites < 0) bad("sites must be non-negative");
  if (auto* v = b.get("spin_twice"); v && !v->is_null()) s.spin_twice = as_int(*v, "spin_twice");
  const Value* up = b.get("up");
  const Value* ups = b.get("up_sectors");
  if (up && !up->is_null() && ups && !ups->is_null()) bad("spin basis accepts either up or up_sectors, not both");
  if (up && !up->is_null()) s.up = as_int(*up, "up");
  if (ups && !ups->is_null()) {
    for (auto& e : ups->array()) s.up_sectors.push_back(as_int(e, "up_sectors"));
    // canonical_particle_sectors: sorted unique, within capacity
    std::sort(s.up_sectors.begin(), s.up_sectors.end());
    s.up_sectors.erase(std::unique(s.up_sectors.begin(), s.up_sectors.end()), s.up_sectors.end());
    if (s.up_sectors.empty()) throw invalid_sector("spin particle sectors must not be empty");
  }
  if (auto* v = b.get("momentum"); v && !v->is_null()) s.momentum = as_int(*v, "momentum");
  if (auto* v = b.get("parity"); v && !v->is_null()) s.parity = as_int(*v, "parity");
  const Value* norm = b.get("normalization");
  if (norm && !norm->is_null()) {
    if (norm->kind != Value::String) bad("normalization must be a string");
    if (norm->raw == "angular_momentum") s.normalization = Normalization::AngularMomentum;
    else if (norm->raw == "pauli") s.normalization = Normalization::Pauli;
    else if (norm->raw == "pauli_cartesian") s.normalization = Normalization::PauliCartesian;
    else bad("invalid binding request: unknown variant `" + norm->raw + "`");
  } else {
    s.normalization = as_bool(b.get("pauli"), false) ? Normalization::PauliCartesian : Normalization::AngularMomentum;
  }
  // SymmetryRequest {destinations, local_permutations?, sector} (capi
  // :906-911); checked against the lattice when the sector is built
  const Value* sym = b.get("symmetries");
  if (sym && !sym->is_null()) {
    for (auto& q : sym->array()) {
      if (q.kind != Value::Obj) bad("invalid binding request: symmetries must hold objects");
      SymmetryMapSpec m;
      for (auto& d : need(q, "destinations", "symmetry").array()) {
        const uint64_t u = as_u64(d, "destinations");
        m.destinations.push_back(u > uint64_t(LONG_MAX) ? LONG_MAX : long(u));
      }
      if (auto* lp = q.get("local_permutations"); lp && !lp->is_null()) {
        for (auto& perm : lp->array()) {
          std::vector<long> digits;
          for (auto& d : perm.array()) {
            const uint64_t u = as_u64(d, "local_permutations");
            digits.push_back(u > uint64_t(LONG_MAX) ? LONG_MAX : long(u));
          }
          m.local_permutations.push_back(std::move(digits));
        }
      }
      m.sector = as_int(need(q, "sector", "symmetry"), "sector");
      s.symmetries.push_back(std::move(m));
    }
  }
  if (!s.symmetries.empty() && (s.momentum || s.parity)) bad("built-in and general spin symmetries cannot be mixed");
  if (auto* v = b.get("matrix_symmetry"); v && !v->is_null())
    bad("matrix-valued symmetry sectors materialize as projected operator models");
  s.reverse = as_bool(b.get("reverse"), false);
  return s;
}

// TermRequest (capi :724-742, local names :6053-6071)
std::vector<OperatorSpec> parse_terms(const Value& t) {
  if (t.kind != Value::Arr) bad("invalid binding request: terms must be an array");
  std::vector<OperatorSpec> out;
  for (auto& term : *t.arr) {
    const Value& prod = need(term, "product", "term");
    const Value& local = need(prod, "local", "product");
    OperatorSpec spec;
    for (auto& name : local.array()) {
      if (name.kind != Value::String) bad("local operator names must be strings");
      const std::string& n = name.raw;
      if (n == "identity") spec.symbols += 'I';
      else if (n == "number") spec.symbols += 'n';
      else if (n == "z") spec.symbols += 'z';
      else if (n == "raising") spec.symbols += '+';
      else if (n == "lowering") spec.symbols += '-';
      else if (n == "x") spec.symbols += 'x';
      else if (n == "y") spec.symbols += 'y';
      else throw invalid_operator(n);
    }
    for (auto& cpl : need(term, "couplings", "term").array()) {
      Coupling c;
      c.coefficient = as_complex(need(cpl, "coefficient", "coupling"), "coefficient");
      for (auto& s : need(cpl, "sites", "coupling").array()) c.sites.push_back(as_int(s, "sites"));
      spec.couplings.push_back(c);
    }
    out.push_back(spec);
  }
  validate_terms(out, 0);
  return out;
}

AssemblyChecks parse_checks(const Value* v) {
  AssemblyChecks c;
  if (!v || v->is_null()) return c;
  c.hermiticity = as_bool(v->get("hermiticity"), true);
  c.particle_conservation = as_bool(v->get("particle_conservation"), true);
  c.symmetry_compatibility = as_bool(v->get("symmetry_compatibility"), true);
  return c;
}

MatrixFormat parse_format(const Value* v) {
  if (!v || v->is_null()) return MatrixFormat::Csc;
  if (v->kind != Value::String) bad("format must be a string");
  const std::string& f = v->raw;
  if (f == "dense") return MatrixFormat::Dense;
  if (f == "csc") return MatrixFormat::Csc;
  if (f == "csr") return MatrixFormat::Csr;
  if (f == "dia") return MatrixFormat::Dia;
  if (f == "matrix_free") return MatrixFormat::MatrixFree;
  bad("invalid binding request: unknown variant `" + f + "`");
}

struct SolverReq {
  EigshOptions opts;
  bool eigenvectors = false;
  std::vector<cplx> initial;
  bool has_initial = false;
  // extension: a whole initial subspace (eigsh_with_initial_subspace,
  // src/solve/mod.rs:2385-2410)
  std::vector<std::vector<cplx>> subspace;
  bool has_subspace = false;
  EigshControls ctl;
};

// SolverRequest (capi :2064-2150): defaults tol 1e-10, iterations 1000, seed 0
SolverReq parse_solver(const Value& s) {
  SolverReq r;
  r.opts.eigenpairs = as_int(need(s, "eigenpairs", "solver"), "eigenpairs");
  const Value& tg = need(s, "target", "solver");
  const Value* kind = tg.get("kind");
  if (!kind || kind->kind != Value::String) bad("invalid binding request: missing field `kind`");
  const std::string& k = kind->raw;
  if (k == "smallest_algebraic") r.opts.target = SpectrumTarget::SmallestAlgebraic;
  else if (k == "largest_algebraic") r.opts.target = SpectrumTarget::LargestAlgebraic;
  else if (k == "smallest_magnitude") r.opts.target = SpectrumTarget::SmallestMagnitude;
  else if (k == "largest_magnitude") r.opts.target = SpectrumTarget::LargestMagnitude;
  else if (k == "both_ends") r.opts.target = SpectrumTarget::BothEnds;
  else if (k == "shift") {
    r.opts.target = SpectrumTarget::Shift;
    r.opts.shift = as_double(need(tg, "value", "target"), "value");
  } else bad("invalid binding request: unknown variant `" + k + "`");
  if (auto* v = s.get("krylov_dimension"); v && !v->is_null()) r.opts.krylov_dimension = as_long(*v, "krylov_dimension");
  if (auto* v = s.get("tolerance"); v && !v->is_null()) r.opts.tolerance = as_double(*v, "tolerance");
  if (auto* v = s.get("max_iterations"); v && !v->is_null()) r.opts.max_iterations = as_long(*v, "max_iterations");
  if (auto* v = s.get("seed"); v && !v->is_null()) r.opts.seed = as_u64(*v, "seed");
  r.eigenvectors = as_bool(s.get("eigenvectors"), false);
  const Value* iv = s.get("initial_vector");
  if (!iv || iv->is_null()) iv = s.get("v0");
  if (iv && !iv->is_null()) {
    r.initial = as_cvec(*iv, "initial_vector");
    r.has_initial = true;
  }
  if (auto* v = s.get("initial_subspace"); v && !v->is_null()) {
    if (r.has_initial) bad("initial_vector and initial_subspace cannot be combined");
    if (v->kind != Value::Arr) bad("initial_subspace must be an array of vectors");
    for (auto& e : *v->arr) r.subspace.push_back(as_cvec(e, "initial_subspace"));
    r.has_subspace = true;
  }
  // engine controls (extensions; absent in reference requests)
  if (auto* v = s.get("matvec_precision"); v && v->kind == Value::String) {
    if (v->raw == "fp32") r.ctl.precision = MatvecPrecision::FP32;
    else if (v->raw == "fp64") r.ctl.precision = MatvecPrecision::FP64;
    else bad("matvec_precision must be \"fp64\" or \"fp32\"");
  }
  if (auto* v = s.get("early_stop"); v && v->kind == Value::Bool) r.ctl.early_stop = v->b;
  if (auto* v = s.get("refill"); v && !v->is_null()) r.ctl.lockstep_refill = as_bool(v, false) ? 1 : 0;
  if (auto* v = s.get("order_by_gap"); v && !v->is_null()) r.ctl.lockstep_order = as_bool(v, true);
  if (auto* v = s.get("refill_group"); v && !v->is_null()) {
    const long g = as_long(*v, "refill_group");
    if (g < 0 || g > 16) bad("refill_group must lie in [0, 16]");
    r.ctl.lockstep_refill = int(g);
  }
  if (auto* v = s.get("batch_rows"); v && !v->is_null()) {
    const long c = as_long(*v, "batch_rows");
    if (c < 0) bad("batch_rows must not be negative");
    r.ctl.batch_rows = c;
  }
  if (auto* v = s.get("check_every"); v && !v->is_null()) {
    const long c = as_long(*v, "check_every");
    if (c < 1 || c > 64) bad("check_every must lie in [1, 64]");
    r.ctl.check_every = int(c);
  }
  if (auto* v = s.get("lockstep"); v && !v->is_null()) {
    const long c = as_long(*v, "lockstep");
    if (c < 0 || c > 16) bad("lockstep must lie in [0, 16]");
    r.ctl.lockstep = int(c);
  }
  if (auto* v = s.get("read_lag"); v && !v->is_null()) {
    const long c = as_long(*v, "read_lag");
    if (c < 0 || c > 8) bad("read_lag must lie in [0, 8]");
    r.ctl.read_lag = int(c);
  }
  if (auto* v = s.get("reorthogonalization_eta"); v && !v->is_null()) {
    const double eta = as_double(*v, "reorthogonalization_eta");
    if (!(eta > 0.0 && eta <= 1.0)) bad("reorthogonalization_eta must lie in (0, 1]");
    r.ctl.dgks_eta = eta;
  }
  return r;
}

// ------------------------------------------------------------------ model --
struct Model {
  SpinBasisSpec spec;
  std::vector<OperatorSpec> terms;
  std::vector<ComponentSpec> components;
  AssemblyChecks checks;
  std::shared_ptr<SpinBasis1D> basis;
  std::shared_ptr<Operator> op;
  hipStream_t stream = nullptr;
  std::mutex mu;
  ~Model() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  Operator& materialized() {
    if (!op)
      op = OperatorBuilder::on(basis).terms(terms).components(components).checks(checks).build(
          MatrixFormat::MatrixFree);
    return *op;
  }
  // Evaluate a family at the request's parameters (caller holds mu). A model
  // without components accepts no parameters (src/interop.rs:964-973).
  Operator& bind(const Value* parameters) {
    Operator& o = materialized();
    std::vector<std::pair<std::string, cplx>> named;
    if (parameters && !parameters->is_null()) {
      if (parameters->kind != Value::Obj) throw invalid_options("parameters must be an object");
      for (auto& kv : *parameters->obj) named.emplace_back(kv.first, as_complex(kv.second, "parameters"));
    }
    if (!o.parameterized()) {
      if (!named.empty()) throw invalid_options("unknown operator parameter " + debug_string(named[0].first));
      return o;
    }
    o.evaluate_coefficients(o.resolve_coefficients(named), stream);
    // PackedOperatorModel checks the evaluated operator (src/interop.rs:589-640)
    if (!o.hermitian()) throw non_hermitian();
    return o;
  }
};

// TermComponentRequest (capi :783-787)
std::vector<ComponentSpec> parse_components(const Value* v) {
  std::vector<ComponentSpec> out;
  if (!v || v->is_null()) return out;
  for (auto& c : v->array()) {
    if (c.kind != Value::Obj) bad("components: expected objects");
    ComponentSpec cs;
    const Value& name = need(c, "name", "component");
    if (name.kind != Value::String) bad("component name must be a string");
    cs.name = name.raw;
    cs.terms = parse_terms(need(c, "terms", "component"));
    if (auto* d = c.get("default"); d && !d->is_null()) {
      cs.has_default = true;
      cs.default_value = as_complex(*d, "default");
    }
    out.push_back(std::move(cs));
  }
  // validate_term_components (src/interop.rs:1308-1326), at model creation
  std::set<std::string> names;
  for (const auto& c : out) {
    if (c.name.empty() || !names.insert(c.name).second)
      throw invalid_options("component names must be nonempty and unique");
    if (c.has_default && (!std::isfinite(c.default_value.real()) || !std::isfinite(c.default_value.imag())))
      throw invalid_options("component defaults must be finite");
  }
  return out;
}

// site_permutation (EdModel::with_site_permutation, src/interop.rs:682-708,
// 1328-1345): every entry below the permutation's length, no repeats
std::vector<int> parse_site_permutation(const Value& p) {
  std::vector<uint64_t> raw;
  for (auto& e : p.array()) raw.push_back(as_u64(e, "site_permutation"));
  const uint64_t len = raw.size();
  for (uint64_t site : raw)
    if (site >= len) throw invalid_site(long(std::min<uint64_t>(site, uint64_t(LONG_MAX))), long(len));
  std::set<uint64_t> uniq(raw.begin(), raw.end());
  if (uniq.size() != raw.size()) bad("site permutation must be bijective");
  return std::vector<int>(raw.begin(), raw.end());
}

// every coupling's site s becomes permutation[s] (OperatorSpec::
// with_site_permutation, src/operator/mod.rs:269-309)
void permute_sites(std::vector<OperatorSpec>& terms, const std::vector<int>& perm) {
  for (auto& t : terms)
    for (auto& c : t.couplings)
      for (int& site : c.sites) {
        if (site < 0 || size_t(site) >= perm.size()) throw invalid_site(site, long(perm.size()));
        site = perm[size_t(site)];
      }
}

std::shared_ptr<Model> make_model(const Value& basis, const Value& terms, const Value* checks,
                                  const Value* components = nullptr, const Value* site_permutation = nullptr) {
  auto m = std::make_shared<Model>();
  m->spec = parse_basis(basis);
  m->terms = parse_terms(terms);
  m->components = parse_components(components);
  if (site_permutation && !site_permutation->is_null()) {
    const std::vector<int> perm = parse_site_permutation(*site_permutation);
    permute_sites(m->terms, perm);
    for (auto& c : m->components) permute_sites(c.terms, perm);
  }
  m->checks = parse_checks(checks);
  QMBED_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  m->basis = SpinBasis1D::build(m->spec, m->stream);
  return m;
}

std::mutex g_reg_mu;
std::map<uint64_t, std::shared_ptr<Model>> g_models;
std::atomic<uint64_t> g_next_model{1};

std::string register_model_json(const std::shared_ptr<Model>& m) {
  const uint64_t id = g_next_model.fetch_add(1);
  const size_t dim = m->basis->len();
  {
    std::lock_guard<std::mutex> g(g_reg_mu);
    g_models[id] = m;
  }
  return "{\"kind\":\"model\",\"handle\":\"" + std::to_string(id) + "\",\"dimension\":" + std::to_string(dim) + "}";
}

// deferred basis plans (create_basis_plan): the request parts, parsed again
// when materialized
struct BasisPlan {
  Value basis;
  Value site_permutation;
  Value checks;
};
std::map<uint64_t, std::shared_ptr<BasisPlan>> g_plans;
std::atomic<uint64_t> g_next_plan{1};

// parse_basis_plan_handle (capi :1706-1718)
uint64_t parse_plan_handle(const std::string& h) {
  char* end = nullptr;
  errno = 0;
  unsigned long long v = h.empty() ? 0 : std::strtoull(h.c_str(), &end, 10);
  if (h.empty() || !end || *end || errno || h[0] == '-' || h[0] == '+') {
    std::string q;
    json::write_string(q, h);
    bad("basis-plan handle " + q + " is not a positive integer");
  }
  if (v == 0) bad("basis-plan handle must be positive");
  return v;
}

std::shared_ptr<BasisPlan> registered_plan(const Value& h) {
  if (h.kind != Value::String) bad("plan_handle must be a string");
  const uint64_t id = parse_plan_handle(h.raw);
  std::lock_guard<std::mutex> g(g_reg_mu);
  auto it = g_plans.find(id);
  if (it == g_plans.end()) bad("basis-plan handle " + std::to_string(id) + " is not registered");
  return it->second;
}

struct ExpmEntry {
  std::shared_ptr<Model> model;
  std::shared_ptr<ExpmActionPlan> plan;
};
std::map<uint64_t, std::shared_ptr<ExpmEntry>> g_expm;
std::atomic<uint64_t> g_next_expm{1};

uint64_t parse_model_handle(const std::string& h) {
  char* end = nullptr;
  errno = 0;
  unsigned long long v = h.empty() ? 0 : std::strtoull(h.c_str(), &end, 10);
  if (h.empty() || !end || *end || errno || h[0] == '-' || h[0] == '+') {
    std::string q;
    json::write_string(q, h);
    bad("model handle " + q + " is not a positive integer");
  }
  if (v == 0) bad("model handle must be positive");
  return v;
}

std::shared_ptr<Model> registered_model(const std::string& h) {
  uint64_t id = parse_model_handle(h);
  std::lock_guard<std::mutex> g(g_reg_mu);
  auto it = g_models.find(id);
  if (it == g_models.end()) bad("model handle " + std::to_string(id) + " is not registered");
  return it->second;
}

uint64_t parse_expm_handle(const std::string& h) {
  std::string q;
  json::write_string(q, h);
  if (h.rfind("expm:", 0) != 0) bad("exponential-action handle " + q + " does not use the expm: prefix");
  std::string rest = h.substr(5);
  char* end = nullptr;
  unsigned long long v = rest.empty() ? 0 : std::strtoull(rest.c_str(), &end, 10);
  if (rest.empty() || !end || *end || rest[0] == '-' || rest[0] == '+') bad("exponential-action handle " + q + " is invalid");
  if (v == 0) bad("exponential-action handle must be positive");
  return v;
}

// ---------------------------------------------------------------- writers --
void write_vec(std::string& o, const std::vector<cplx>& v) {
  o += '[';
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) o += ',';
    o += '[';
    json::write_number(o, v[i].real());
    o += ',';
    json::write_number(o, v[i].imag());
    o += ']';
  }
  o += ']';
}

void write_dvec(std::string& o, const std::vector<double>& v) {
  o += '[';
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) o += ',';
    json::write_number(o, v[i]);
  }
  o += ']';
}

std::string eigensystem_body(size_t dim, const Eigensystem& es, bool vectors, bool tagged) {
  std::string o = "{";
  if (tagged) o += "\"kind\":\"eigensystem\",";
  o += "\"dimension\":" + std::to_string(dim) + ",\"eigenvalues\":";
  write_dvec(o, es.eigenvalues);
  o += ",\"residuals\":";
  write_dvec(o, es.residuals);
  o += ",\"iterations\":" + std::to_string(es.iterations);
  o += std::string(",\"converged\":") + (es.converged ? "true" : "false");
  if (vectors) {
    o += ",\"eigenvectors\":[";
    for (size_t q = 0; q < es.eigenvectors.size(); ++q) {
      if (q) o += ',';
      write_vec(o, es.eigenvectors[q]);
    }
    o += "]";
  }
  o += "}";
  return o;
}

Eigensystem solve_model(Model& m, const SolverReq& r, const Value* parameters = nullptr,
                        std::optional<MatrixFormat> format = std::nullopt) {
  std::lock_guard<std::mutex> g(m.mu);
  Operator& op = m.bind(parameters);
  if (format) op.use_format(*format, m.stream);
  if (r.has_initial && r.initial.size() != op.dim())
    throw dimension_mismatch("eigsh initial vector does not match the operator");
  if (r.has_subspace) return eigsh_subspace(op, r.opts, r.subspace, r.ctl, r.eigenvectors, m.stream);
  return eigsh(op, r.opts, r.has_initial ? r.initial.data() : nullptr, r.ctl, r.eigenvectors, m.stream);
}

std::string counters_body(size_t dim, const Eigensystem& es, bool vectors) {
  std::string body = eigensystem_body(dim, es, vectors, false);
  body.pop_back();
  body += ",\"fp32_iterations\":" + std::to_string(es.fp32_iterations) +
          ",\"fp64_iterations\":" + std::to_string(es.fp64_iterations) +
          ",\"reorthogonalization_passes\":" + std::to_string(es.reorthogonalization_passes) +
          ",\"conditional_second_passes\":" + std::to_string(es.conditional_second_passes) + "}";
  return body;
}

// Many models in one lockstep batch (eigsh_sectors). Each model is used at
// its current evaluation; the batch runs on the first model's stream.
std::string solve_models(const std::vector<std::shared_ptr<Model>>& models, const SolverReq& r) {
  if (models.empty()) bad("the batch names no models");
  if (r.has_initial || r.has_subspace) bad("the sector batch starts every sector from the deterministic start vector");
  std::vector<Model*> order;
  for (auto& m : models) order.push_back(m.get());
  std::sort(order.begin(), order.end());
  if (std::adjacent_find(order.begin(), order.end()) != order.end()) bad("a model appears twice in the batch");
  std::vector<std::unique_lock<std::mutex>> locks;
  for (Model* m : order) locks.emplace_back(m->mu);
  std::vector<Operator*> ops;
  for (auto& m : models) ops.push_back(&m->materialized());
  SectorBatchStats st;
  auto res = eigsh_sectors(ops, r.opts, r.ctl, r.eigenvectors, models[0]->stream, &st);
  std::string o = "{\"kind\":\"eigensystems\",\"results\":[";
  for (size_t i = 0; i < res.size(); ++i) {
    if (i) o += ',';
    o += counters_body(ops[i]->dim(), res[i], r.eigenvectors);
  }
  o += "],\"batch\":{\"sectors\":" + std::to_string(st.sectors) + ",\"batched\":" + std::to_string(st.batched) +
       ",\"fallback\":" + std::to_string(st.fallback) + ",\"steps\":" + std::to_string(st.steps) +
       ",\"ritz_rounds\":" + std::to_string(st.ritz_rounds) + ",\"window\":" + std::to_string(st.window) +
       ",\"batches\":" + std::to_string(st.batches) +
       ",\"seconds\":";
  json::write_number(o, st.seconds);
  o += "}}";
  return o;
}

// y_c = A x_c (conj_io: conj(A conj(x_c))) for host columns, through the
// model's current action
std::vector<std::vector<cplx>> apply_columns(const Operator& op, const std::vector<std::vector<cplx>>& cols,
                                             bool conj_io, hipStream_t s) {
  const size_t n = op.dim();
  std::vector<std::vector<cplx>> out;
  if (cols.empty()) return out;
  DeviceArray<cplx> x(std::max<size_t>(n, 1)), y(std::max<size_t>(n, 1));
  for (const auto& c : cols) {
    std::vector<cplx> in = c;
    if (conj_io)
      for (auto& e : in) e = std::conj(e);
    std::vector<cplx> res(n);
    if (n) {
      QMBED_HIP(hipMemcpyAsync(x.data(), in.data(), n * sizeof(cplx), hipMemcpyHostToDevice, s));
      op.apply(x.data(), y.data(), s);
      QMBED_HIP(hipMemcpyAsync(res.data(), y.data(), n * sizeof(cplx), hipMemcpyDeviceToHost, s));
      QMBED_HIP(hipStreamSynchronize(s));
    }
    if (conj_io)
      for (auto& e : res) e = std::conj(e);
    out.push_back(std::move(res));
  }
  return out;
}

std::string run_solve(const std::string& req) {
  Value v;
  try {
    v = json::parse(req);
  } catch (json::ParseError& e) {
    throw invalid_options(std::string("invalid binding request: ") + e.what());
  }
  if (v.kind != Value::Obj) bad("invalid binding request: expected an object");
  auto m = make_model(need(v, "basis", "invalid binding request"), need(v, "terms", "invalid binding request"),
                      nullptr);
  const MatrixFormat format = parse_format(v.get("format"));
  SolverReq r = parse_solver(need(v, "solver", "invalid binding request"));
  Eigensystem es = solve_model(*m, r, nullptr, format);
  return "{\"status\":\"ok\",\"result\":" + eigensystem_body(m->basis->len(), es, r.eigenvectors, false) + "}";
}

std::string ok(const std::string& body) { return "{\"status\":\"ok\",\"result\":" + body + "}"; }

std::string run_command(const std::string& req) {
  Value v;
  try {
    v = json::parse(req);
  } catch (json::ParseError& e) {
    throw invalid_options(std::string("invalid binding command: ") + e.what());
  }
  const Value* opv = v.get("operation");
  if (!opv || opv->kind != Value::String) bad("invalid binding command: missing field `operation`");
  const std::string& op = opv->raw;

  if (op == "describe_basis") {
    SpinBasisSpec spec = parse_basis(need(v, "basis", "describe_basis"));
    auto b = SpinBasis1D::build(spec, nullptr);
    auto states = b->row_states_host64();
    std::string o = "{\"kind\":\"basis\",\"dimension\":" + std::to_string(b->len()) + ",\"states\":[";
    for (size_t i = 0; i < states.size(); ++i) {
      if (i) o += ',';
      o += '"' + std::to_string(states[i]) + '"';
    }
    o += "]}";
    return ok(o);
  }
  if (op == "create_model") {
    auto m = make_model(need(v, "basis", "create_model"), need(v, "terms", "create_model"), v.get("checks"),
                        v.get("components"), v.get("site_permutation"));
    return ok(register_model_json(m));
  }
  if (op == "create_basis_plan") {
    // create_basis_plan (capi :3753-3768): the request kept for a later
    // materialization; the basis is checked now
    auto plan = std::make_shared<BasisPlan>();
    plan->basis = need(v, "basis", "create_basis_plan");
    if (auto* p = v.get("site_permutation"); p && !p->is_null()) plan->site_permutation = *p;
    if (auto* c = v.get("checks"); c && !c->is_null()) plan->checks = *c;
    (void)parse_basis(plan->basis);
    (void)parse_checks(v.get("checks"));
    if (auto* p = v.get("site_permutation"); p && !p->is_null()) (void)parse_site_permutation(*p);
    const uint64_t id = g_next_plan.fetch_add(1);
    {
      std::lock_guard<std::mutex> g(g_reg_mu);
      g_plans[id] = plan;
    }
    return ok("{\"kind\":\"basis_plan\",\"handle\":\"" + std::to_string(id) + "\"}");
  }
  if (op == "materialize_basis_plan") {
    // materialize_basis_plan (capi :3770-3781): a model with no terms on the
    // plan's basis, checks and site permutation
    auto plan = registered_plan(need(v, "plan_handle", "materialize_basis_plan"));
    json::Value empty;
    empty.kind = json::Value::Arr;
    empty.arr = std::make_shared<std::vector<json::Value>>();
    auto m = make_model(plan->basis, empty, plan->checks.kind == json::Value::Null ? nullptr : &plan->checks, nullptr,
                        plan->site_permutation.kind == json::Value::Null ? nullptr : &plan->site_permutation);
    return ok(register_model_json(m));
  }
  if (op == "release_basis_plan") {
    const Value& h = need(v, "plan_handle", "release_basis_plan");
    if (h.kind != Value::String) bad("plan_handle must be a string");
    const uint64_t id = parse_plan_handle(h.raw);
    std::lock_guard<std::mutex> g(g_reg_mu);
    if (!g_plans.erase(id)) bad("basis-plan handle " + std::to_string(id) + " is not registered");
    return ok("{\"kind\":\"released\",\"handle\":\"" + std::to_string(id) + "\"}");
  }

  if (op == "eigsh_model") {
    const Value& h = need(v, "handle", "eigsh_model");
    if (h.kind != Value::String) bad("handle must be a string");
    auto m = registered_model(h.raw);
    const MatrixFormat format = parse_format(v.get("format"));
    SolverReq r = parse_solver(need(v, "solver", "eigsh_model"));
    Eigensystem es = solve_model(*m, r, v.get("parameters"), format);
    return ok(eigensystem_body(m->basis->len(), es, r.eigenvectors, true));
  }
  if (op == "describe_model") {
    // registered_describe_model (capi :5990-5997): the sector's rows
    const Value& h = need(v, "handle", "describe_model");
    if (h.kind != Value::String) bad("handle must be a string");
    auto m = registered_model(h.raw);
    auto states = m->basis->row_states_host64();
    std::string o = "{\"kind\":\"basis\",\"dimension\":" + std::to_string(m->basis->len()) + ",\"states\":[";
    for (size_t i = 0; i < states.size(); ++i) {
      if (i) o += ',';
      o += '"' + std::to_string(states[i]) + '"';
    }
    o += "]}";
    return ok(o);
  }
  if (op == "materialize_model") {
    // command_operator (capi :3913-3940): the stored entries in row-major
    // order for the requested format (Operator::triplets)
    const Value& h = need(v, "handle", "materialize_model");
    if (h.kind != Value::String) bad("handle must be a string");
    auto m = registered_model(h.raw);
    const MatrixFormat format = parse_format(v.get("format"));
    std::vector<Operator::Triplet> trip;
    size_t n = 0;
    {
      std::lock_guard<std::mutex> g(m->mu);
      Operator& o = m->bind(v.get("parameters"));
      n = o.dim();
      trip = o.triplets(m->stream);
    }
    if (format == MatrixFormat::Dia) {
      std::set<long> diagonals;
      for (auto& t : trip) diagonals.insert(long(t.row) - long(t.col));
      if (diagonals.size() > std::max<size_t>(2 * n, 1))
        throw unsupported_backend("the requested operator is not usefully diagonal-banded");
    }
    static const char* names[] = {"dense", "csc", "csr", "dia", "matrix_free"};
    std::string o = "{\"kind\":\"operator\",\"shape\":[" + std::to_string(n) + "," + std::to_string(n) +
                    "],\"format\":\"" + names[int(format)] + "\",\"entries\":[";
    for (size_t i = 0; i < trip.size(); ++i) {
      if (i) o += ',';
      o += "{\"row\":" + std::to_string(trip[i].row) + ",\"column\":" + std::to_string(trip[i].col) + ",\"value\":[";
      json::write_number(o, trip[i].value.real());
      o += ',';
      json::write_number(o, trip[i].value.imag());
      o += "]}";
    }
    o += "]}";
    return ok(o);
  }
  if (op == "apply_model") {
    // command_apply_model (capi :4351-4360): A·x for each vector, with the
    // requested action (normal, transpose, conjugate, adjoint)
    const Value& h = need(v, "handle", "apply_model");
    if (h.kind != Value::String) bad("handle must be a string");
    auto m = registered_model(h.raw);
    std::string action = "normal";
    if (auto* a = v.get("action"); a && !a->is_null()) {
      if (a->kind != Value::String) bad("action must be a string");
      action = a->raw;
      if (action != "normal" && action != "transpose" && action != "conjugate" && action != "adjoint")
        bad("invalid binding command: unknown variant `" + action + "`");
    }
    std::vector<std::vector<cplx>> cols;
    for (auto& c : need(v, "vectors", "apply_model").array()) cols.push_back(as_cvec(c, "vectors"));
    std::vector<std::vector<cplx>> out;
    size_t n = 0;
    {
      std::lock_guard<std::mutex> g(m->mu);
      Operator& o = m->bind(v.get("parameters"));
      n = o.dim();
      for (auto& c : cols)
        if (c.size() != n) throw dimension_mismatch("operator input has the wrong length");
      // a Hermitian matrix is its own adjoint, and its transpose is its
      // conjugate: A^T x = conj(A conj(x))
      const bool herm = o.hermitian();
      const bool conj_io = action == "conjugate" || (herm && action == "transpose");
      if (!herm && (action == "transpose" || action == "adjoint"))
        throw unsupported_backend("transposed actions of a non-Hermitian sector matrix are not built on the MI355X path");
      out = apply_columns(o, cols, conj_io, m->stream);
    }
    std::string o = "{\"kind\":\"vectors\",\"dimension\":" + std::to_string(n) + ",\"vectors\":[";
    for (size_t c = 0; c < out.size(); ++c) {
      if (c) o += ',';
      write_vec(o, out[c]);
    }
    o += "]}";
    return ok(o);
  }
  if (op == "eigh_model") {
    // command_eigh (capi :5290-5299): the whole spectrum of the dense matrix
    const Value& h = need(v, "handle", "eigh_model");
    if (h.kind != Value::String) bad("handle must be a string");
    auto m = registered_model(h.raw);
    const bool vectors = as_bool(v.get("eigenvectors"), false);
    Eigensystem es;
    {
      std::lock_guard<std::mutex> g(m->mu);
      Operator& o = m->bind(v.get("parameters"));
      if (!o.hermitian()) throw non_hermitian();
      es = eigh_all(o, vectors, m->stream);
    }
    return ok(eigensystem_body(m->basis->len(), es, vectors, true));
  }
  if (op == "eigsh_models") {
    // extension: several registered models (sectors) solved in one batch
    std::vector<std::shared_ptr<Model>> ms;
    for (auto& h : need(v, "handles", "eigsh_models").array()) {
      if (h.kind != Value::String) bad("handle must be a string");
      ms.push_back(registered_model(h.raw));
    }
    SolverReq r = parse_solver(need(v, "solver", "eigsh_models"));
    return ok(solve_models(ms, r));
  }
  if (op == "evolve_model") {
    const Value& h = need(v, "handle", "evolve_model");
    if (h.kind != Value::String) bad("handle must be a string");
    auto m = registered_model(h.raw);
    const Value& ev = need(v, "evolution", "evolve_model");
    EvolutionOptions eo;
    for (auto& t : need(ev, "times", "evolution").array()) eo.times.push_back(as_double(t, "times"));
    eo.krylov_dimension = 64;  // capi default (:596-606)
    if (auto* x = ev.get("krylov_dimension"); x && !x->is_null()) eo.krylov_dimension = as_long(*x, "krylov_dimension");
    if (auto* x = ev.get("tolerance"); x && !x->is_null()) eo.tolerance = as_double(*x, "tolerance");
    if (auto* x = ev.get("max_substeps"); x && !x->is_null()) eo.max_substeps = as_long(*x, "max_substeps");
    if (as_bool(ev.get("imaginary_time"), false))
      bad("static imaginary-time evolution is not exposed by this command yet");
    std::vector<std::vector<cplx>> cols;
    for (auto& c : need(v, "vectors", "evolve_model").array()) cols.push_back(as_cvec(c, "vectors"));
    std::lock_guard<std::mutex> g(m->mu);
    const Operator& H = m->bind(v.get("parameters"));
    auto traj = evolve_batch(H, cols, eo, nullptr, m->stream);
    std::string o = "{\"kind\":\"trajectory\",\"dimension\":" + std::to_string(H.dim()) + ",\"times\":";
    write_dvec(o, eo.times);
    o += ",\"states\":[";
    for (size_t t = 0; t < traj.size(); ++t) {
      if (t) o += ',';
      o += '[';
      for (size_t c = 0; c < traj[t].size(); ++c) {
        if (c) o += ',';
        write_vec(o, traj[t][c]);
      }
      o += ']';
    }
    o += "]}";
    return ok(o);
  }
  if (op == "create_expm_action") {
    const Value& h = need(v, "handle", "create_expm_action");
    if (h.kind != Value::String) bad("handle must be a string");
    auto m = registered_model(h.raw);
    cplx coef = as_complex(need(v, "coefficient", "create_expm_action"), "coefficient");
    long max_degree = 55;
    double tol = 0.5 * 2.220446049250313e-16;
    long max_sub = 10000;
    if (auto* x = v.get("max_degree"); x && !x->is_null()) max_degree = as_long(*x, "max_degree");
    if (auto* x = v.get("tolerance"); x && !x->is_null()) tol = as_double(*x, "tolerance");
    if (auto* x = v.get("max_substeps"); x && !x->is_null()) max_sub = as_long(*x, "max_substeps");
    {
      std::lock_guard<std::mutex> g(m->mu);
      if (m->materialized().parameterized())
        throw unsupported_backend("exponential-action plans of parameterized models are not built on the MI355X path");
      m->bind(v.get("parameters"));
    }
    std::shared_ptr<ExpmEntry> e = std::make_shared<ExpmEntry>();
    e->model = m;
    {
      std::lock_guard<std::mutex> g(m->mu);
      m->materialized();
      e->plan = std::make_shared<ExpmActionPlan>(m->op, coef, int(max_degree), tol, max_sub, m->stream);
    }
    uint64_t id = g_next_expm.fetch_add(1);
    {
      std::lock_guard<std::mutex> g(g_reg_mu);
      g_expm[id] = e;
    }
    return ok("{\"kind\":\"expm_plan\",\"handle\":\"expm:" + std::to_string(id) +
              "\",\"dimension\":" + std::to_string(m->basis->len()) + "}");
  }
  if (op == "apply_expm_action") {
    const Value& h = need(v, "expm_handle", "apply_expm_action");
    if (h.kind != Value::String) bad("expm_handle must be a string");
    uint64_t id = parse_expm_handle(h.raw);
    std::shared_ptr<ExpmEntry> e;
    {
      std::lock_guard<std::mutex> g(g_reg_mu);
      auto it = g_expm.find(id);
      if (it == g_expm.end()) bad("exponential-action handle " + h.raw + " is not registered");
      e = it->second;
    }
    if (auto* t = v.get("threads"); t && !t->is_null()) {
      long th = as_long(*t, "threads");
      if (th <= 0) bad("runtime thread count must be positive");
    }
    std::vector<std::vector<cplx>> cols;
    for (auto& c : need(v, "vectors", "apply_expm_action").array()) cols.push_back(as_cvec(c, "vectors"));
    for (auto& c : cols)
      if (c.size() != e->plan->dimension()) throw dimension_mismatch("exponential batch column has the wrong length");
    std::vector<std::vector<cplx>> out;
    {
      std::lock_guard<std::mutex> g(e->model->mu);
      out = e->plan->apply(cols, e->model->stream);
    }
    size_t dim = out.empty() ? e->plan->dimension() : out[0].size();
    std::string o = "{\"kind\":\"vectors\",\"dimension\":" + std::to_string(dim) + ",\"vectors\":[";
    for (size_t c = 0; c < out.size(); ++c) {
      if (c) o += ',';
      write_vec(o, out[c]);
    }
    o += "]}";
    return ok(o);
  }
  if (op == "release_model") {
    const Value& h = need(v, "handle", "release_model");
    if (h.kind != Value::String) bad("handle must be a string");
    uint64_t id = parse_model_handle(h.raw);
    std::lock_guard<std::mutex> g(g_reg_mu);
    if (!g_models.erase(id)) bad("model handle " + std::to_string(id) + " is not registered");
    return ok("{\"kind\":\"released\",\"handle\":\"" + std::to_string(id) + "\"}");
  }
  if (op == "release_expm_action") {
    const Value& h = need(v, "expm_handle", "release_expm_action");
    if (h.kind != Value::String) bad("expm_handle must be a string");
    uint64_t id = parse_expm_handle(h.raw);
    std::lock_guard<std::mutex> g(g_reg_mu);
    if (!g_expm.erase(id)) bad("exponential-action handle " + h.raw + " is not registered");
    return ok("{\"kind\":\"released\",\"handle\":\"expm:" + std::to_string(id) + "\"}");
  }
  throw unsupported_backend("operation `" + op + "` is outside the MI355X spin eigsh/evolve/expm path");
}

// The first invalid UTF-8 sequence of a NUL-terminated request, described
// as the reference's binding reports it (the display of Rust's Utf8Error):
// "invalid utf-8 sequence of {n} bytes from index {i}" for a bad sequence,
// "incomplete utf-8 byte sequence from index {i}" for one cut short by the
// end. Well-formed UTF-8 (RFC 3629: no overlong forms, no surrogates, at
// most U+10FFFF) gives nothing.
std::optional<std::string> utf8_error(const char* text) {
  const auto* v = reinterpret_cast<const unsigned char*>(text);
  const size_t len = std::strlen(text);
  auto cont = [](unsigned char c) { return (c & 0xC0u) == 0x80u; };
  size_t i = 0;
  while (i < len) {
    const unsigned char c = v[i];
    if (c < 0x80) {
      ++i;
      continue;
    }
    const size_t start = i;
    const int width = c >= 0xC2 && c <= 0xDF ? 2 : c >= 0xE0 && c <= 0xEF ? 3 : c >= 0xF0 && c <= 0xF4 ? 4 : 0;
    auto invalid = [&](int n) {
      return "invalid utf-8 sequence of " + std::to_string(n) + " bytes from index " + std::to_string(start);
    };
    auto incomplete = [&]() { return "incomplete utf-8 byte sequence from index " + std::to_string(start); };
    if (width == 0) return invalid(1);
    if (i + 1 >= len) return incomplete();
    const unsigned char c1 = v[i + 1];
    bool ok1;
    if (width == 2) {
      ok1 = cont(c1);
    } else if (width == 3) {
      ok1 = (c == 0xE0 && c1 >= 0xA0 && c1 <= 0xBF) || (c >= 0xE1 && c <= 0xEC && cont(c1)) ||
            (c == 0xED && c1 >= 0x80 && c1 <= 0x9F) || (c >= 0xEE && cont(c1));
    } else {
      ok1 = (c == 0xF0 && c1 >= 0x90 && c1 <= 0xBF) || (c >= 0xF1 && c <= 0xF3 && cont(c1)) ||
            (c == 0xF4 && c1 >= 0x80 && c1 <= 0x8F);
    }
    if (!ok1) return invalid(1);
    for (int k = 2; k < width; ++k) {
      if (i + size_t(k) >= len) return incomplete();
      if (!cont(v[i + size_t(k)])) return invalid(k);
    }
    i += size_t(width);
  }
  return std::nullopt;
}

template <typename F>
char* guarded(const char* request, F&& f) {
  std::string resp;
  try {
    if (!request) {
      resp = "{\"status\":\"error\",\"error\":\"request pointer is null\"}";
    } else if (auto bad_utf8 = utf8_error(request)) {
      resp = error_json("request is not UTF-8: " + *bad_utf8);
    } else {
      resp = f(std::string(request));
    }
  } catch (Error& e) {
    resp = error_json(e.message);
  } catch (std::exception& e) {
    resp = error_json(std::string("internal state error: ") + e.what());
  } catch (...) {
    resp = error_json("internal state error: native binding failure");
  }
  return dup_c(resp);
}

// the argument checks of the reference's registration entry points
// (bindings/capi/src/lib.rs:6207-6216, :6253-6262), then the answer that user
// bases are not built here
template <typename Op, typename Map>
char* user_basis_unsupported(const char* request, Op op, const Map* maps, size_t map_count) {
  if (request && !op) return dup_c("{\"status\":\"error\",\"error\":\"user operator callback is null\"}");
  if (request && map_count > 0 && !maps)
    return dup_c("{\"status\":\"error\",\"error\":\"user symmetry callback array is null\"}");
  return guarded(request, [](const std::string&) -> std::string {
    throw unsupported_backend("user-defined bases are outside the MI355X spin-sector path");
  });
}

}  // namespace

// ------------------------------------------------------------------ typed --
struct TypedModel {
  std::shared_ptr<Model> m;
};

struct TypedScan {
  std::shared_ptr<Model> m;
  std::unique_ptr<WarmStartScan> scan;
};

json::Value parse_or_throw(const std::string& req) {
  try {
    return json::parse(req);
  } catch (json::ParseError& e) {
    throw invalid_options(std::string("invalid binding request: ") + e.what());
  }
}

void write_long_array(std::string& o, const std::vector<long>& v) {
  o += '[';
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) o += ',';
    o += std::to_string(v[i]);
  }
  o += ']';
}

}  // namespace qmbed

using namespace qmbed;

extern "C" {

char* qmbed_run_json(const char* request) { return guarded(request, run_solve); }

char* qmbed_command_json(const char* request) { return guarded(request, run_command); }

char* qmbed_evolve_model_with_drive_json(const char* request, QmbedDriveCallback callback, void* context) {
  (void)context;
  if (request && !callback) return dup_c("{\"status\":\"error\",\"error\":\"drive callback is null\"}");
  return guarded(request, [](const std::string&) -> std::string {
    throw unsupported_backend("driven (time-dependent) evolution is outside the MI355X static-Hamiltonian path");
  });
}

char* qmbed_register_user_basis_32_json(const char* request, QmbedUserOp32Callback op,
                                        QmbedUserNextState32Callback next_state,
                                        QmbedUserPreCheck32Callback pre_check,
                                        const QmbedUserMap32Callback* map_callbacks, size_t map_count) {
  (void)next_state;
  (void)pre_check;
  return user_basis_unsupported(request, op, map_callbacks, map_count);
}

char* qmbed_register_user_basis_64_json(const char* request, QmbedUserOp64Callback op,
                                        QmbedUserNextState64Callback next_state,
                                        QmbedUserPreCheck64Callback pre_check,
                                        const QmbedUserMap64Callback* map_callbacks, size_t map_count) {
  (void)next_state;
  (void)pre_check;
  return user_basis_unsupported(request, op, map_callbacks, map_count);
}

void qmbed_string_free(char* s) { std::free(s); }

// ----------------------------------------------------------- typed path ----
qmbed_hip_model* qmbed_hip_model_create(const char* model_json, char** error) {
  try {
    if (!model_json) throw invalid_options("request pointer is null");
    json::Value v;
    try {
      v = json::parse(model_json);
    } catch (json::ParseError& e) {
      throw invalid_options(std::string("invalid binding request: ") + e.what());
    }
    auto tm = new TypedModel();
    try {
      tm->m = make_model(need(v, "basis", "model"), need(v, "terms", "model"), v.get("checks"), v.get("components"));
      Operator& op = tm->m->materialized();
      // an optional stored format ("csr", "csc", "dense", "dia"): the typed
      // products then run on the stored matrix, as the JSON requests do
      if (const Value* f = v.get("format"); f && !f->is_null()) op.use_format(parse_format(f), tm->m->stream);
    } catch (...) {
      delete tm;
      throw;
    }
    if (error) *error = nullptr;
    return reinterpret_cast<qmbed_hip_model*>(tm);
  } catch (Error& e) {
    if (error) *error = dup_c(e.message);
  } catch (std::exception& e) {
    if (error) *error = dup_c(std::string("internal state error: ") + e.what());
  }
  return nullptr;
}

void qmbed_hip_model_free(qmbed_hip_model* m) { delete reinterpret_cast<TypedModel*>(m); }

long qmbed_hip_model_dimension(const qmbed_hip_model* m) {
  return long(reinterpret_cast<const TypedModel*>(m)->m->basis->len());
}

int qmbed_hip_model_is_real(const qmbed_hip_model* m) {
  return reinterpret_cast<const TypedModel*>(m)->m->op->is_real() ? 1 : 0;
}

int qmbed_hip_model_states(const qmbed_hip_model* m, uint32_t* states, uint8_t* orbits) {
  try {
    auto& b = *reinterpret_cast<const TypedModel*>(m)->m->basis;
    if (b.wide()) return -2;
    auto s = b.row_states_host();
    std::memcpy(states, s.data(), s.size() * 4);
    if (orbits) {
      auto o = b.row_orbits_host();
      std::memcpy(orbits, o.data(), o.size());
    }
    return 0;
  } catch (...) {
    return -1;
  }
}

int qmbed_hip_model_states64(const qmbed_hip_model* m, uint64_t* states, uint8_t* orbits) {
  try {
    auto& b = *reinterpret_cast<const TypedModel*>(m)->m->basis;
    auto s = b.row_states_host64();
    std::memcpy(states, s.data(), s.size() * 8);
    if (orbits) {
      auto o = b.row_orbits_host();
      std::memcpy(orbits, o.data(), o.size());
    }
    return 0;
  } catch (...) {
    return -1;
  }
}

long qmbed_hip_model_nnz(const qmbed_hip_model* m) {
  try {
    auto* tm = reinterpret_cast<const TypedModel*>(m);
    return long(tm->m->op->stored_entries(tm->m->stream));
  } catch (...) {
    return -1;
  }
}

int qmbed_hip_model_hv_config(const qmbed_hip_model* m, int* lanes_per_row, int* image_table) {
  try {
    auto* tm = reinterpret_cast<const TypedModel*>(m);
    if (lanes_per_row) *lanes_per_row = tm->m->op->lanes_per_row();
    if (image_table) *image_table = tm->m->op->uses_index_stream() ? 2 : (tm->m->op->uses_image_table() ? 1 : 0);
    return 0;
  } catch (...) {
    return -1;
  }
}

void* qmbed_hip_model_stream(const qmbed_hip_model* m) { return reinterpret_cast<const TypedModel*>(m)->m->stream; }

int qmbed_hip_apply_device(const qmbed_hip_model* m, int dtype, const void* x, void* y, void* stream) {
  try {
    auto* tm = reinterpret_cast<const TypedModel*>(m);
    const Operator& op = *tm->m->op;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : tm->m->stream;
    switch (dtype) {
      case QMBED_F64: op.apply(static_cast<const double*>(x), static_cast<double*>(y), s); break;
      case QMBED_F32: op.apply(static_cast<const float*>(x), static_cast<float*>(y), s); break;
      case QMBED_C128: op.apply(static_cast<const cplx*>(x), static_cast<cplx*>(y), s); break;
      case QMBED_C64:
        op.apply(static_cast<const std::complex<float>*>(x), static_cast<std::complex<float>*>(y), s);
        break;
      case QMBED_F32_TO_F64:
      case QMBED_C64_TO_C128:
        if (!op.apply_mixed(static_cast<const float*>(x), static_cast<double*>(y), dtype == QMBED_C64_TO_C128, s))
          return -3;
        break;
      default: return -2;
    }
    return 0;
  } catch (...) {
    return -1;
  }
}

// element bytes of the input (out = false) or output vector
static size_t dtype_size(int dtype, bool out = false) {
  switch (dtype) {
    case QMBED_F64: return 8;
    case QMBED_F32: return 4;
    case QMBED_C128: return 16;
    case QMBED_C64: return 8;
    case QMBED_F32_TO_F64: return out ? 8 : 4;
    case QMBED_C64_TO_C128: return out ? 16 : 8;
    default: return 0;
  }
}

int qmbed_hip_apply_host(const qmbed_hip_model* m, int dtype, const void* x, void* y) {
  try {
    auto* tm = reinterpret_cast<const TypedModel*>(m);
    size_t n = tm->m->basis->len();
    size_t bytes = dtype_size(dtype) * n, out_bytes = dtype_size(dtype, true) * n;
    if (!bytes) return dtype_size(dtype) ? 0 : -2;
    DeviceArray<uint8_t> dx(bytes), dy(out_bytes);
    QMBED_HIP(hipMemcpy(dx.data(), x, bytes, hipMemcpyHostToDevice));
    int rc = qmbed_hip_apply_device(m, dtype, dx.data(), dy.data(), nullptr);
    if (rc) return rc;
    QMBED_HIP(hipStreamSynchronize(tm->m->stream));
    QMBED_HIP(hipMemcpy(y, dy.data(), out_bytes, hipMemcpyDeviceToHost));
    return 0;
  } catch (...) {
    return -1;
  }
}

double qmbed_hip_time_apply(const qmbed_hip_model* m, int dtype, const void* x, void* y, int iters) {
  try {
    auto* tm = reinterpret_cast<const TypedModel*>(m);
    hipStream_t s = tm->m->stream;
    hipEvent_t a, b;
    QMBED_HIP(hipEventCreate(&a));
    QMBED_HIP(hipEventCreate(&b));
    QMBED_HIP(hipEventRecord(a, s));
    for (int i = 0; i < iters; ++i)
      if (qmbed_hip_apply_device(m, dtype, x, y, s) != 0) return -1.0;
    QMBED_HIP(hipEventRecord(b, s));
    QMBED_HIP(hipEventSynchronize(b));
    float ms = 0;
    QMBED_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return double(ms) / double(std::max(iters, 1));
  } catch (...) {
    return -1.0;
  }
}

double qmbed_hip_time_scan_batch(qmbed_hip_model* m, int width, int fp32, const double* points, int iters) {
  try {
    auto* tm = reinterpret_cast<TypedModel*>(m);
    std::lock_guard<std::mutex> g(tm->m->mu);
    Operator& op = tm->m->materialized();
    hipStream_t s = tm->m->stream;
    if (!(width == 1 || width == 2 || width == 4 || width == 8 || width == 16)) return -2.0;
    if (width > op.scan_batch_max_width() || !op.scan_batch_supported(s) || !points || iters < 1) return -2.0;
    const size_t n = op.dim();
    const int count = op.scan_batch_amp_count();
    const size_t P = tm->m->components.size();
    std::vector<std::vector<cplx>> coefs(static_cast<size_t>(width));
    for (int b = 0; b < width; ++b)
      for (size_t p = 0; p < P; ++p) coefs[size_t(b)].push_back(cplx(points[size_t(b) * P + p], 0.0));
    DeviceArray<double> amps(size_t(width) * size_t(count)), diags(size_t(width) * 2 * n), ys(size_t(width) * n);
    const size_t xbytes = (fp32 ? 4 : 8) * size_t(width) * n;
    DeviceArray<uint8_t> x(xbytes);
    op.scan_batch_tables(coefs, amps.data(), diags.data(), s);
    // every lane holds the deterministic start vector
    DeviceArray<double> v(n);
    deterministic_start_device(v.data(), n, 0, s, false);
    std::vector<const double*> src(static_cast<size_t>(width), v.data());
    lk::interleave(src.data(), n, width, fp32 ? nullptr : reinterpret_cast<double*>(x.data()),
                   fp32 ? reinterpret_cast<float*>(x.data()) : nullptr, s);
    const double* am[16] = {};
    const double* dg[16] = {};
    double* yp[16] = {};
    for (int b = 0; b < width; ++b) {
      am[b] = amps.data() + size_t(b) * count;
      dg[b] = diags.data() + size_t(b) * 2 * n;
      yp[b] = ys.data() + size_t(b) * n;
    }
    for (int i = 0; i < 2; ++i) op.apply_scan_batch(x.data(), fp32 != 0, width, am, dg, yp, nullptr, s);
    hipEvent_t a, b;
    QMBED_HIP(hipEventCreate(&a));
    QMBED_HIP(hipEventCreate(&b));
    QMBED_HIP(hipEventRecord(a, s));
    for (int i = 0; i < iters; ++i) op.apply_scan_batch(x.data(), fp32 != 0, width, am, dg, yp, nullptr, s);
    QMBED_HIP(hipEventRecord(b, s));
    QMBED_HIP(hipEventSynchronize(b));
    float ms = 0;
    QMBED_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return double(ms) / double(iters);
  } catch (...) {
    return -1.0;
  }
}

int qmbed_hip_apply_scan_batch(qmbed_hip_model* m, int width, int fp32, const double* points, const double* x,
                               double* y) {
  try {
    auto* tm = reinterpret_cast<TypedModel*>(m);
    std::lock_guard<std::mutex> g(tm->m->mu);
    Operator& op = tm->m->materialized();
    hipStream_t s = tm->m->stream;
    if (!(width == 1 || width == 2 || width == 4 || width == 8 || width == 16)) return -2;
    if (width > op.scan_batch_max_width() || !op.scan_batch_supported(s)) return -2;
    if (!points || !x || !y) return -1;
    const size_t n = op.dim();
    if (n == 0) return 0;
    const int count = op.scan_batch_amp_count();
    const size_t P = tm->m->components.size();
    std::vector<std::vector<cplx>> coefs(static_cast<size_t>(width));
    for (int b = 0; b < width; ++b)
      for (size_t p = 0; p < P; ++p) coefs[size_t(b)].push_back(cplx(points[size_t(b) * P + p], 0.0));
    DeviceArray<double> amps(size_t(width) * size_t(count)), diags(size_t(width) * 2 * n), ys(size_t(width) * n);
    DeviceArray<double> cols(size_t(width) * n);
    DeviceArray<uint8_t> xi((fp32 ? 4 : 8) * size_t(width) * n);
    op.scan_batch_tables(coefs, amps.data(), diags.data(), s);
    QMBED_HIP(hipMemcpyAsync(cols.data(), x, size_t(width) * n * 8, hipMemcpyHostToDevice, s));
    std::vector<const double*> src(static_cast<size_t>(width));
    for (int b = 0; b < width; ++b) src[size_t(b)] = cols.data() + size_t(b) * n;
    lk::interleave(src.data(), n, width, fp32 ? nullptr : reinterpret_cast<double*>(xi.data()),
                   fp32 ? reinterpret_cast<float*>(xi.data()) : nullptr, s);
    const double* am[16] = {};
    const double* dg[16] = {};
    double* yp[16] = {};
    for (int b = 0; b < width; ++b) {
      am[b] = amps.data() + size_t(b) * count;
      dg[b] = diags.data() + size_t(b) * 2 * n;
      yp[b] = ys.data() + size_t(b) * n;
    }
    op.apply_scan_batch(xi.data(), fp32 != 0, width, am, dg, yp, nullptr, s);
    QMBED_HIP(hipMemcpyAsync(y, ys.data(), size_t(width) * n * 8, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
    return 0;
  } catch (...) {
    return -1;
  }
}

int qmbed_hip_launch_timer(int on) {
  try {
    launch_timer_switch(on != 0);
    return 0;
  } catch (...) {
    return -1;
  }
}

char* qmbed_hip_launch_timer_report(void) {
  try {
    return dup_c("{\"status\":\"ok\",\"result\":" + launch_timer_report() + "}");
  } catch (Error& e) {
    return dup_c(error_json(e.message));
  } catch (std::exception& e) {
    return dup_c(error_json(std::string("internal state error: ") + e.what()));
  }
}

char* qmbed_hip_eigsh_json(const qmbed_hip_model* m, const char* solver_json) {
  return guarded(solver_json, [&](const std::string& req) {
    auto* tm = reinterpret_cast<const TypedModel*>(m);
    json::Value v;
    try {
      v = json::parse(req);
    } catch (json::ParseError& e) {
      throw invalid_options(std::string("invalid binding request: ") + e.what());
    }
    SolverReq r = parse_solver(v);
    Eigensystem es = solve_model(*tm->m, r, v.get("parameters"));
    return ok(counters_body(tm->m->basis->len(), es, r.eigenvectors));
  });
}

qmbed_hip_workspace* qmbed_hip_workspace_create(void) {
  return reinterpret_cast<qmbed_hip_workspace*>(new EigshWorkspace());
}

void qmbed_hip_workspace_free(qmbed_hip_workspace* w) { delete reinterpret_cast<EigshWorkspace*>(w); }

long qmbed_hip_workspace_dimension(const qmbed_hip_workspace* w) {
  return long(reinterpret_cast<const EigshWorkspace*>(w)->dimension);
}

int qmbed_hip_workspace_vectors(const qmbed_hip_workspace* w) {
  return int(reinterpret_cast<const EigshWorkspace*>(w)->subspace.size());
}

char* qmbed_hip_workspace_set(qmbed_hip_workspace* w, long dimension, int count, const double* vectors) {
  return guarded("{}", [&](const std::string&) {
    if (dimension < 0 || count < 0 || (count > 0 && !vectors)) bad("workspace vectors must be given");
    std::vector<std::vector<cplx>> vs(static_cast<size_t>(count));
    for (int j = 0; j < count; ++j) {
      const double* v = vectors + size_t(j) * 2 * size_t(dimension);
      vs[size_t(j)].resize(size_t(dimension));
      for (long i = 0; i < dimension; ++i) vs[size_t(j)][size_t(i)] = cplx(v[2 * i], v[2 * i + 1]);
    }
    reinterpret_cast<EigshWorkspace*>(w)->set(size_t(dimension), std::move(vs));
    return ok("{\"kind\":\"workspace\",\"dimension\":" + std::to_string(dimension) +
              ",\"vectors\":" + std::to_string(count) + "}");
  });
}

char* qmbed_hip_eigsh_workspace_json(const qmbed_hip_model* m, qmbed_hip_workspace* w, const char* solver_json) {
  return guarded(solver_json, [&](const std::string& req) {
    auto* tm = reinterpret_cast<const TypedModel*>(m);
    auto& ws = *reinterpret_cast<EigshWorkspace*>(w);
    json::Value v = parse_or_throw(req);
    SolverReq r = parse_solver(v);
    if (r.has_initial || r.has_subspace) bad("a workspace solve takes its start from the workspace");
    std::lock_guard<std::mutex> g(tm->m->mu);
    const Operator& op = tm->m->bind(v.get("parameters"));
    Eigensystem es = eigsh_workspace(op, r.opts, ws, r.ctl, tm->m->stream);
    return ok(counters_body(tm->m->basis->len(), es, r.eigenvectors));
  });
}

char* qmbed_hip_evolve_device(qmbed_hip_model* m, const void* states, int ncols, void* out,
                              const char* evolution_json) {
  return guarded(evolution_json, [&](const std::string& req) {
    auto* tm = reinterpret_cast<TypedModel*>(m);
    json::Value v;
    try {
      v = json::parse(req);
    } catch (json::ParseError& e) {
      throw invalid_options(std::string("invalid binding request: ") + e.what());
    }
    if (ncols < 0) bad("ncols must not be negative");
    if (ncols > 0 && (!states || !out)) bad("states and out must be device pointers");
    EvolutionOptions eo;  // EvolutionOptions::new defaults (Krylov 30)
    for (auto& t : need(v, "times", "evolution").array()) eo.times.push_back(as_double(t, "times"));
    if (auto* x = v.get("krylov_dimension"); x && !x->is_null()) eo.krylov_dimension = as_long(*x, "krylov_dimension");
    if (auto* x = v.get("tolerance"); x && !x->is_null()) eo.tolerance = as_double(*x, "tolerance");
    if (auto* x = v.get("max_substeps"); x && !x->is_null()) eo.max_substeps = as_long(*x, "max_substeps");
    std::lock_guard<std::mutex> g(tm->m->mu);
    const Operator& H = tm->m->bind(v.get("parameters"));
    EvolutionDiagnostics dg;
    evolve_batch_device(H, static_cast<const cplx*>(states), ncols, eo, &dg, static_cast<cplx*>(out), tm->m->stream);
    std::string o = "{\"kind\":\"evolution\",\"dimension\":" + std::to_string(H.dim()) + ",\"times\":";
    write_dvec(o, eo.times);
    o += ",\"lanczos_projections\":" + std::to_string(dg.lanczos_projections) +
         ",\"matrix_vector_products\":" + std::to_string(dg.matrix_vector_products) +
         ",\"real_lanczos_projections\":" + std::to_string(dg.real_lanczos_projections) +
         ",\"real_matrix_vector_products\":" + std::to_string(dg.real_matrix_vector_products) +
         ",\"accepted_substeps\":" + std::to_string(dg.accepted_substeps) +
         ",\"rejected_trial_intervals\":" + std::to_string(dg.rejected_trial_intervals) +
         ",\"maximum_estimated_error\":";
    json::write_number(o, dg.maximum_estimated_error);
    o += "}";
    return ok(o);
  });
}

char* qmbed_hip_eigsh_batch_json(qmbed_hip_model* const* models, int count, const char* solver_json) {
  return guarded(solver_json, [&](const std::string& req) {
    if (count <= 0 || !models) bad("the batch names no models");
    std::vector<std::shared_ptr<Model>> ms;
    for (int i = 0; i < count; ++i) {
      if (!models[i]) bad("the batch holds a null model");
      ms.push_back(reinterpret_cast<TypedModel*>(models[i])->m);
    }
    json::Value v;
    try {
      v = json::parse(req);
    } catch (json::ParseError& e) {
      throw invalid_options(std::string("invalid binding request: ") + e.what());
    }
    return ok(solve_models(ms, parse_solver(v)));
  });
}

char* qmbed_hip_model_set_parameters(qmbed_hip_model* m, const char* parameters_json) {
  return guarded(parameters_json, [&](const std::string& req) {
    auto* tm = reinterpret_cast<TypedModel*>(m);
    json::Value v;
    try {
      v = json::parse(req);
    } catch (json::ParseError& e) {
      throw invalid_options(std::string("invalid binding request: ") + e.what());
    }
    std::lock_guard<std::mutex> g(tm->m->mu);
    Operator& op = tm->m->bind(&v);
    std::string body = "{\"kind\":\"coefficients\",\"values\":[";
    const auto& c = op.coefficients();
    for (size_t i = 0; i < c.size(); ++i) {
      if (i) body += ',';
      body += '[';
      json::write_number(body, c[i].real());
      body += ',';
      json::write_number(body, c[i].imag());
      body += ']';
    }
    body += "],\"real\":" + std::string(op.is_real() ? "true" : "false") + "}";
    return ok(body);
  });
}

qmbed_hip_scan* qmbed_hip_scan_create(qmbed_hip_model* model, const char* options_json, char** error) {
  try {
    if (!model) throw invalid_options("model pointer is null");
    auto* tm = reinterpret_cast<TypedModel*>(model);
    ScanOptions o;
    if (options_json) {
      json::Value v = parse_or_throw(options_json);
      if (v.kind != json::Value::Obj) bad("scan options must be an object");
      if (auto* x = v.get("snapshots"); x && !x->is_null()) o.snapshots = as_int(*x, "snapshots");
      if (auto* x = v.get("hidden"); x && !x->is_null()) {
        auto& a = x->array();
        if (a.size() != 2) bad("hidden must hold two layer widths");
        o.hidden1 = as_int(a[0], "hidden");
        o.hidden2 = as_int(a[1], "hidden");
      }
      if (auto* x = v.get("epochs"); x && !x->is_null()) o.epochs = as_int(*x, "epochs");
      if (auto* x = v.get("learning_rate"); x && !x->is_null()) o.learning_rate = float(as_double(*x, "learning_rate"));
      o.final_learning_rate = o.learning_rate;  // constant unless given
      if (auto* x = v.get("final_learning_rate"); x && !x->is_null())
        o.final_learning_rate = float(as_double(*x, "final_learning_rate"));
      if (auto* x = v.get("seed"); x && !x->is_null()) o.seed = as_u64(*x, "seed");
      if (auto* x = v.get("targets"); x && !x->is_null()) {
        if (x->kind != json::Value::String) bad("targets must be a string");
        if (x->raw == "projections") o.targets = ScanOptions::Targets::Projections;
        else if (x->raw == "continuation") o.targets = ScanOptions::Targets::Continuation;
        else if (x->raw == "correction") o.targets = ScanOptions::Targets::Correction;
        else bad("targets must be \"projections\", \"continuation\" or \"correction\"");
      }
      if (auto* x = v.get("provisional"); x && x->kind == json::Value::Bool) o.provisional = x->b;
      if (auto* x = v.get("space"); x && !x->is_null()) {
        if (x->kind != json::Value::String) bad("space must be a string");
        if (x->raw == "greedy") o.space = ScanOptions::Space::Greedy;
        else if (x->raw == "principal") o.space = ScanOptions::Space::Principal;
        else bad("space must be \"greedy\" or \"principal\"");
      }
      if (auto* x = v.get("samples"); x && !x->is_null()) {
        o.samples = as_int(*x, "samples");
        if (o.samples < 0 || o.samples > 65536) bad("samples must lie in [0, 65536]");
      }
    }
    auto ts = std::make_unique<TypedScan>();
    ts->m = tm->m;
    std::lock_guard<std::mutex> g(ts->m->mu);
    ts->scan = std::make_unique<WarmStartScan>(ts->m->materialized(), ts->m->stream, o);
    if (error) *error = nullptr;
    return reinterpret_cast<qmbed_hip_scan*>(ts.release());
  } catch (Error& e) {
    if (error) *error = dup_c(e.message);
  } catch (std::exception& e) {
    if (error) *error = dup_c(std::string("internal state error: ") + e.what());
  }
  return nullptr;
}

void qmbed_hip_scan_free(qmbed_hip_scan* scan) { delete reinterpret_cast<TypedScan*>(scan); }

char* qmbed_hip_scan_train(qmbed_hip_scan* scan, const double* points, int n, const char* solver_json) {
  return guarded(solver_json, [&](const std::string& req) {
    auto* ts = reinterpret_cast<TypedScan*>(scan);
    if (!points) throw invalid_options("points pointer is null");
    SolverReq r = parse_solver(parse_or_throw(req));
    std::lock_guard<std::mutex> g(ts->m->mu);
    auto rep = ts->scan->train(points, n, r.opts, r.ctl);
    std::string o = "{\"kind\":\"scan_training\",\"points\":" + std::to_string(n) +
                    ",\"snapshots\":" + std::to_string(rep.snapshots) + ",\"energies\":";
    write_dvec(o, rep.energies);
    o += ",\"iterations\":";
    write_long_array(o, rep.iterations);
    o += ",\"provisional_starts\":" + std::to_string(rep.provisional_starts);
    o += ",\"snapshot_indices\":";
    write_long_array(o, std::vector<long>(rep.snapshot_indices.begin(), rep.snapshot_indices.end()));
    o += ",\"solve_seconds\":";
    json::write_number(o, rep.solve_seconds);
    o += ",\"gram_seconds\":";
    json::write_number(o, rep.gram_seconds);
    o += ",\"train_seconds\":";
    json::write_number(o, rep.train_seconds);
    o += ",\"loss_first\":";
    json::write_number(o, rep.loss_first);
    o += ",\"loss_last\":";
    json::write_number(o, rep.loss_last);
    o += ",\"target_norm2\":";
    json::write_number(o, rep.target_norm2);
    o += "}";
    return ok(o);
  });
}

char* qmbed_hip_scan_solve(qmbed_hip_scan* scan, const double* points, int n, const char* solver_json, int warm) {
  return guarded(solver_json, [&](const std::string& req) {
    auto* ts = reinterpret_cast<TypedScan*>(scan);
    if (!points) throw invalid_options("points pointer is null");
    SolverReq r = parse_solver(parse_or_throw(req));
    std::lock_guard<std::mutex> g(ts->m->mu);
    std::vector<double> e, res, secs;
    std::vector<long> it, it32, it64;
    long warm_n = 0, fallbacks = 0;
    const auto start = warm == 2 ? WarmStartScan::Start::SnapshotRR
                                 : (warm ? WarmStartScan::Start::Network : WarmStartScan::Start::Plain);
    for (const auto& pr : ts->scan->solve_many(points, n, r.opts, r.ctl, start)) {
      e.push_back(pr.es.eigenvalues[0]);
      res.push_back(pr.es.residuals[0]);
      secs.push_back(pr.seconds);
      it.push_back(pr.es.iterations);
      it32.push_back(pr.es.fp32_iterations);
      it64.push_back(pr.es.fp64_iterations);
      warm_n += pr.warm ? 1 : 0;
      fallbacks += pr.fallback ? 1 : 0;
    }
    std::string o = "{\"kind\":\"scan_solve\",\"points\":" + std::to_string(n) + ",\"energies\":";
    write_dvec(o, e);
    o += ",\"residuals\":";
    write_dvec(o, res);
    o += ",\"seconds\":";
    write_dvec(o, secs);
    o += ",\"iterations\":";
    write_long_array(o, it);
    o += ",\"fp32_iterations\":";
    write_long_array(o, it32);
    o += ",\"fp64_iterations\":";
    write_long_array(o, it64);
    o += ",\"warm_starts\":" + std::to_string(warm_n) + ",\"fallbacks\":" + std::to_string(fallbacks) + "}";
    return ok(o);
  });
}

int qmbed_hip_scan_predict(qmbed_hip_scan* scan, const double* points, int n, double* coefficients) {
  try {
    auto* ts = reinterpret_cast<TypedScan*>(scan);
    std::lock_guard<std::mutex> g(ts->m->mu);
    auto y = ts->scan->predict(points, n);
    std::memcpy(coefficients, y.data(), y.size() * sizeof(double));
    return ts->scan->snapshots();
  } catch (...) {
    return -1;
  }
}

int qmbed_hip_mlp_parameter_count(int in, int h1, int h2, int out) {
  try {
    return mlp_parameter_count(in, h1, h2, out);
  } catch (...) {
    return -1;
  }
}

int qmbed_hip_mlp_forward(int in, int h1, int h2, int out, const float* params, const float* x, int n, float* y) {
  try {
    const int np = mlp_parameter_count(in, h1, h2, out);
    DeviceArray<float> dp(static_cast<size_t>(np)), dx(static_cast<size_t>(n) * in), dy(static_cast<size_t>(n) * out);
    QMBED_HIP(hipMemcpy(dp.data(), params, size_t(np) * 4, hipMemcpyHostToDevice));
    QMBED_HIP(hipMemcpy(dx.data(), x, size_t(n) * in * 4, hipMemcpyHostToDevice));
    mlp_forward_device(in, h1, h2, out, dp.data(), dx.data(), n, dy.data(), nullptr);
    QMBED_HIP(hipMemcpy(y, dy.data(), size_t(n) * out * 4, hipMemcpyDeviceToHost));
    return 0;
  } catch (...) {
    return -1;
  }
}

int qmbed_hip_mlp_train(int in, int h1, int h2, int out, float* params, const float* x, const float* t, int n,
                        int epochs, float learning_rate, float* grad, float* loss_history) {
  try {
    const int np = mlp_parameter_count(in, h1, h2, out);
    DeviceArray<float> dp(static_cast<size_t>(np)), dx(static_cast<size_t>(n) * in), dt(static_cast<size_t>(n) * out),
        dg(static_cast<size_t>(np)), dl(1), dh(static_cast<size_t>(std::max(epochs, 1)));
    QMBED_HIP(hipMemcpy(dp.data(), params, size_t(np) * 4, hipMemcpyHostToDevice));
    QMBED_HIP(hipMemcpy(dx.data(), x, size_t(n) * in * 4, hipMemcpyHostToDevice));
    QMBED_HIP(hipMemcpy(dt.data(), t, size_t(n) * out * 4, hipMemcpyHostToDevice));
    mlp_train_device(in, h1, h2, out, dp.data(), dx.data(), dt.data(), n, epochs, learning_rate, 1.0f,
                     epochs == 0 ? dg.data() : nullptr, dl.data(), loss_history ? dh.data() : nullptr, nullptr);
    QMBED_HIP(hipDeviceSynchronize());
    if (epochs == 0) {
      if (grad) QMBED_HIP(hipMemcpy(grad, dg.data(), size_t(np) * 4, hipMemcpyDeviceToHost));
      if (loss_history) QMBED_HIP(hipMemcpy(loss_history, dl.data(), 4, hipMemcpyDeviceToHost));
    } else {
      QMBED_HIP(hipMemcpy(params, dp.data(), size_t(np) * 4, hipMemcpyDeviceToHost));
      if (loss_history) QMBED_HIP(hipMemcpy(loss_history, dh.data(), size_t(epochs) * 4, hipMemcpyDeviceToHost));
    }
    return 0;
  } catch (...) {
    return -1;
  }
}

void* qmbed_hip_malloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}
void qmbed_hip_free(void* p) { (void)hipFree(p); }
int qmbed_hip_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  return hipMemcpy(dst, src, bytes, k) == hipSuccess ? 0 : -1;
}
int qmbed_hip_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }

int qmbed_hip_gemm_atb(const double* A, size_t n, int p, const double* B, int q, double* C) {
  try {
    if (!A || !B || !C || p < 0 || q < 0) return -1;
    DeviceArray<double> part;
    lk::gemm_atb(A, n, p, B, q, C, part, nullptr);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
  } catch (...) {
    return -1;
  }
}
int qmbed_hip_gemm_atb_noalloc(const double* A, size_t n, int p, const double* B, int q, double* C) {
  try {
    if (!A || !B || !C || p < 0 || q < 0) return -1;
    if (!lk::gemm_atb_scratch(A, n, p, B, q, C, nullptr)) return -2;
    return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
  } catch (...) {
    return -1;
  }
}

int qmbed_hip_device_count(void) { return device_count(); }

char* qmbed_hip_basis_archive_save(const char* request, const uint64_t* states, size_t n) {
  try {
    if (!request) bad("invalid binding request: request pointer is null");
    Value v;
    try {
      v = json::parse(request);
    } catch (json::ParseError& e) {
      throw invalid_options(std::string("invalid binding request: ") + e.what());
    }
    if (v.kind != Value::Obj) bad("invalid binding request: expected an object");
    const Value& path = need(v, "path", "invalid binding request");
    if (path.kind != Value::String) bad("invalid binding request: `path` must be a string");
    const long width = as_long(need(v, "width_bits", "invalid binding request"), "width_bits");
    if (width < 0) bad("width_bits: expected a nonnegative integer");
    BasisArchive a;
    if (states) {
      if (v.get("states")) bad("invalid binding request: states come from the request or the buffer, not both");
      a = basis_archive(size_t(width), states, n);
    } else {
      std::vector<std::string> text;
      for (const Value& x : need(v, "states", "invalid binding request").array()) {
        if (x.kind == Value::String) text.push_back(x.raw);
        else if (x.kind == Value::Number) text.push_back(x.raw);  // exact integer text as written
        else bad("invalid binding request: a state must be a decimal string or an integer");
      }
      a = basis_archive(size_t(width), text);
    }
    if (const Value* md = v.get("metadata"); md && !md->is_null()) {
      if (md->kind != Value::Obj) bad("invalid binding request: `metadata` must be an object of strings");
      for (const auto& [k, x] : *md->obj) {
        if (x.kind != Value::String) bad("invalid binding request: metadata values must be strings");
        a.insert_metadata(k, x.raw);
      }
    }
    save_basis_zip(path.raw, a);
    std::string o = "{\"status\":\"ok\",\"result\":{\"path\":";
    json::write_string(o, path.raw);
    o += ",\"width_bits\":" + std::to_string(a.width_bits) + ",\"count\":" + std::to_string(a.states.size()) + "}}";
    return dup_c(o);
  } catch (Error& e) {
    return dup_c(error_json(e.message));
  } catch (std::exception& e) {
    return dup_c(error_json(std::string("internal state error: ") + e.what()));
  }
}

char* qmbed_hip_basis_archive_load(const char* path, uint64_t* states, size_t capacity, int decimal) {
  try {
    if (!path) bad("invalid binding request: path pointer is null");
    const BasisArchive a = load_basis_zip(path);
    bool narrow = true;
    for (const auto& s : a.states)
      if (!fits_u64(s)) {
        narrow = false;
        break;
      }
    const bool written = states && narrow && a.states.size() <= capacity;
    if (written)
      for (size_t i = 0; i < a.states.size(); ++i) states[i] = std::stoull(a.states[i]);
    std::string o = "{\"status\":\"ok\",\"result\":{\"width_bits\":" + std::to_string(a.width_bits) +
                    ",\"count\":" + std::to_string(a.states.size()) + ",\"fits_u64\":" + (narrow ? "true" : "false") +
                    ",\"written\":" + (written ? "true" : "false") + ",\"metadata\":{";
    bool first = true;
    for (const auto& [k, x] : a.metadata) {
      if (!first) o += ",";
      first = false;
      json::write_string(o, k);
      o += ":";
      json::write_string(o, x);
    }
    o += "}";
    if (decimal) {
      o += ",\"states\":[";
      for (size_t i = 0; i < a.states.size(); ++i) {
        if (i) o += ",";
        o += "\"" + a.states[i] + "\"";
      }
      o += "]";
    }
    o += "}}";
    return dup_c(o);
  } catch (Error& e) {
    return dup_c(error_json(e.message));
  } catch (std::exception& e) {
    return dup_c(error_json(std::string("internal state error: ") + e.what()));
  }
}

}  // extern "C"
