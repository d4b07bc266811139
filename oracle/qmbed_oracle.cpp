// QMBED CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A single-threaded FP64 restatement of the reference algorithms on the
// symmetry-reduced spin-chain path: sector enumeration, the column-streaming
// operator assembly into a row-sorted triplet list, the stored matvec, the
// fully reorthogonalized Lanczos eigsh with its dispatch, Krylov evolution with
// the Simpson error bound, and the truncated-Taylor exponential-action plan.
//
// Only the tests, the smoke check and the CPU-baseline leg of the benchmark may
// load this library, and only as the checker. The product library never links
// it and has no CPU fallback.
//
// Each routine names the reference file:line it restates (paths relative to
// the QMBED source tree).
//
// Not restated here: shift-invert targets, matrix-valued (2-D irrep)
// symmetry sectors.

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace oracle {

using cplx = std::complex<double>;
using u64 = std::uint64_t;
static const double EPS = std::numeric_limits<double>::epsilon();

struct OracleError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------- basis ----

// Gosper enumeration of fixed-weight states, ascending.
// src/basis/mod.rs:373-424
static std::vector<u64> fixed_weight_states(int sites, int particles) {
  if (sites > 63) throw OracleError("unsupported backend: oracle states are limited to 63 sites");
  std::vector<u64> states;
  if (particles < 0) {
    u64 limit = u64(1) << sites;
    states.reserve(limit);
    for (u64 s = 0; s < limit; ++s) states.push_back(s);
    return states;
  }
  if (particles > sites) throw OracleError("invalid Hilbert-space sector: particle count exceeds site count");
  if (particles == 0) return {0};
  if (particles == sites) return {(u64(1) << sites) - 1};
  u64 state = (u64(1) << particles) - 1;
  u64 limit = u64(1) << sites;
  for (;;) {
    states.push_back(state);
    u64 low = state & (~state + 1);
    u64 ripple = state + low;
    u64 next = (((ripple ^ state) >> 2) / low) | ripple;
    if (next >= limit) break;
    state = next;
  }
  return states;
}

// src/basis/mod.rs:661-676 (base 2)
static u64 rotate_state(u64 s, int shift, int L) {
  shift %= L;
  if (shift == 0) return s;
  u64 mask = (u64(1) << L) - 1;
  return ((s << shift) & mask) | (s >> (L - shift));
}

// src/basis/mod.rs:688-698 (base 2)
static u64 reflect_state(u64 s, int L) {
  u64 r = 0;
  for (int site = 0; site < L; ++site)
    if ((s >> site) & 1) r |= u64(1) << (L - site - 1);
  return r;
}

struct Image {
  u64 rep;
  cplx phase;
  int orbit;
};

enum Norm { ANGULAR = 0, PAULI = 1, PAULI_CARTESIAN = 2 };

// One generator of a general symmetry sector: destinations[source] is the
// target site, local[source] the permutation of the site's two digits, and
// sector the character index (LatticeSymmetryMap, src/basis/mod.rs:2967-3060)
struct GeneralMap {
  std::vector<int> dest;
  std::vector<std::array<int, 2>> local;
  int sector = 0;
  int period = 1;
};

struct Basis {
  int L = 0;
  int up = -1;                  // -1: no fixed up count
  std::vector<int> up_sectors;  // union of up sectors (sorted), empty: none
  bool has_k = false;
  int k = 0;
  int parity = 0;  // 0: none
  int norm = ANGULAR;
  // general symmetry sector (GeneralBasis + SymmetryReducer) and the
  // descending-row ("reverse") convention
  std::vector<GeneralMap> gens;
  bool reverse = false;
  std::vector<u64> states;
  std::vector<int> orbit;
  std::unordered_map<u64, Image> lookup;
  std::unordered_map<u64, long> rows;  // row of each state when not ascending
  bool general() const { return !gens.empty(); }
  bool symmetric() const { return has_k || parity != 0 || general(); }
  long index(u64 s) const {
    if (reverse || general()) {
      auto it = rows.find(s);
      return it == rows.end() ? -1 : it->second;
    }
    auto it = std::lower_bound(states.begin(), states.end(), s);
    if (it == states.end() || *it != s) return -1;
    return long(it - states.begin());
  }
};

// spin_symmetry_sector: src/basis/mod.rs:715-848
static void symmetry_sector(Basis& b, std::vector<u64> parents) {
  int L = b.L;
  if (b.parity != 0 && b.parity != 1 && b.parity != -1)
    throw OracleError("invalid Hilbert-space sector: parity must be either -1 or +1");
  int kn = 0;
  if (b.has_k) kn = ((b.k % L) + L) % L;
  if (b.parity != 0 && b.has_k && kn != 0 && 2 * kn != L)
    throw OracleError(
        "incompatible symmetry request: parity can share a one-dimensional sector with momentum "
        "only at k=0 or k=pi");
  b.k = kn;
  if (!b.symmetric()) {
    b.states = parents;
    b.orbit.assign(parents.size(), 1);
    return;
  }
  std::unordered_set<u64> parent_set(parents.begin(), parents.end());
  int translations = b.has_k ? L : 1;
  std::unordered_set<u64> visited;
  std::vector<std::pair<u64, int>> sectors;
  const double TAU = 6.283185307179586;
  for (u64 seed : parents) {
    if (visited.count(seed)) continue;
    std::vector<u64> orbit_states;
    for (int s = 0; s < translations; ++s) {
      u64 t = rotate_state(seed, s, L);
      orbit_states.push_back(t);
      if (b.parity) orbit_states.push_back(reflect_state(t, L));
    }
    for (u64 t : orbit_states)
      if (!parent_set.count(t))
        throw OracleError(
            "incompatible symmetry request: symmetry map leaves the selected magnetization sector");
    u64 rep = orbit_states[0];
    for (u64 t : orbit_states) {
      visited.insert(t);
      rep = std::min(rep, t);
    }
    // ordered map keeps iteration deterministic; accumulation order follows
    // the reference's shift loop (src/basis/mod.rs:790-806)
    std::map<u64, cplx> coeff;
    for (int s = 0; s < translations; ++s) {
      double angle = b.has_k ? -TAU * double(kn * s) / double(L) : 0.0;
      cplx ch = std::polar(1.0, angle);
      u64 t = rotate_state(rep, s, L);
      coeff[t] += ch;
      if (b.parity) coeff[reflect_state(t, L)] += double(b.parity) * ch;
    }
    for (auto it = coeff.begin(); it != coeff.end();) {
      if (std::abs(it->second) > 1.0e-12) ++it;
      else it = coeff.erase(it);
    }
    if (coeff.empty()) continue;
    auto rit = coeff.find(rep);
    if (rit == coeff.end())
      throw OracleError(
          "incompatible symmetry request: symmetry projection removed its orbit representative");
    cplx gauge = rit->second / std::abs(rit->second);
    double nrm = 0.0;
    for (auto& kv : coeff) nrm += std::norm(kv.second);
    nrm = std::sqrt(nrm);
    int osize = int(coeff.size());
    double expected = 1.0 / std::sqrt(double(osize));
    for (auto& kv : coeff) {
      cplx normalized = kv.second / (gauge * nrm);
      if (std::abs(std::abs(normalized) - expected) > 1.0e-10)
        throw OracleError(
            "incompatible symmetry request: symmetry projection does not define a one-dimensional "
            "orbit sector");
      b.lookup[kv.first] = Image{rep, normalized / expected, osize};
    }
    sectors.push_back({rep, osize});
  }
  std::sort(sectors.begin(), sectors.end());
  for (auto& s : sectors) {
    b.states.push_back(s.first);
    b.orbit.push_back(s.second);
  }
}

// LatticeSymmetryMap::mapped_state for spin one-half (distinguishable
// sites, so the map phase is 1): src/basis/mod.rs:3100-3132
static u64 apply_map(const GeneralMap& g, u64 s) {
  u64 out = 0;
  for (size_t src = 0; src < g.dest.size(); ++src) {
    const int digit = int((s >> src) & 1);
    out |= u64(g.local[src][size_t(digit)]) << g.dest[src];
  }
  return out;
}

// LatticeSymmetryMap::new checks and combined_permutation_period:
// src/basis/mod.rs:2978-3060, 3221-3268
static void validate_map(GeneralMap& g, int L) {
  if (int(g.dest.size()) != L) {
    std::ostringstream os;
    os << "invalid options: symmetry map has " << g.dest.size() << " sites, expected " << L;
    throw OracleError(os.str());
  }
  if (g.dest.empty()) throw OracleError("invalid Hilbert-space sector: a lattice symmetry requires sites and local states");
  std::vector<int> seen(size_t(L), 0);
  for (int d : g.dest) {
    if (d < 0 || d >= L || seen[size_t(d)]++)
      throw OracleError("incompatible symmetry request: symmetry site destinations must be a bijection");
  }
  for (auto& p : g.local)
    if (!((p[0] == 0 && p[1] == 1) || (p[0] == 1 && p[1] == 0)))
      throw OracleError("incompatible symmetry request: every local-state map must be a permutation");
  // period: lcm of the cycle lengths of (site, digit) -> (dest, local digit)
  const int elements = 2 * L;
  std::vector<char> visited(size_t(elements), 0);
  long period = 1;
  for (int seed = 0; seed < elements; ++seed) {
    if (visited[size_t(seed)]) continue;
    int cur = seed;
    long cycle = 0;
    while (!visited[size_t(cur)]) {
      visited[size_t(cur)] = 1;
      ++cycle;
      const int src = cur / 2, digit = cur % 2;
      cur = g.dest[size_t(src)] * 2 + g.local[size_t(src)][size_t(digit)];
    }
    long a = period, bb = cycle;
    while (bb) {
      long t = a % bb;
      a = bb;
      bb = t;
    }
    period = period / a * cycle;
  }
  g.period = int(period);
}

// GeneralBasis::from_reducer_impl with orbit seeds (no parent-membership
// requirement) and trace_symmetry_orbit: src/basis/mod.rs:3464-3541,
// 4163-4300. Representatives are the minimum or (reverse) maximum orbit
// state; rows follow the parent's (ascending) order of the representatives.
static void general_sector(Basis& b, const std::vector<u64>& parents) {
  const double TAU = 6.283185307179586;
  std::vector<cplx> steps;
  for (auto& g : b.gens) {
    const long nsec = ((long(g.sector) % g.period) + g.period) % g.period;
    steps.push_back(std::polar(1.0, -TAU * double(nsec) / double(g.period)));
  }
  std::unordered_set<u64> visited;
  std::vector<std::pair<u64, int>> reps;
  for (u64 seed : parents) {
    if (visited.count(seed)) continue;
    // closure of every generator at its period (phases are 1 here)
    for (auto& g : b.gens) {
      u64 c = seed;
      for (int i = 0; i < g.period; ++i) c = apply_map(g, c);
      if (c != seed) throw OracleError("incompatible symmetry request: a symmetry map does not close at its declared period");
    }
    std::map<u64, cplx> coeff;  // ordered: deterministic iteration
    coeff[seed] = cplx(1, 0);
    std::deque<u64> queue{seed};
    bool compatible = true;
    while (!queue.empty()) {
      const u64 src = queue.front();
      queue.pop_front();
      const cplx cs = coeff[src];
      for (size_t gi = 0; gi < b.gens.size(); ++gi) {
        const u64 tgt = apply_map(b.gens[gi], src);
        const cplx ct = cs * steps[gi];
        auto it = coeff.find(tgt);
        if (it != coeff.end()) {
          if (std::abs(it->second - ct) > 1.0e-10) compatible = false;
        } else {
          coeff[tgt] = ct;
          queue.push_back(tgt);
        }
      }
    }
    for (auto& kv : coeff) visited.insert(kv.first);
    if (!compatible) continue;
    const u64 rep = b.reverse ? coeff.rbegin()->first : coeff.begin()->first;
    const cplx rc = coeff[rep];
    const cplx gauge = rc / std::abs(rc);
    double nrm = 0;
    for (auto& kv : coeff) nrm += std::norm(kv.second);
    nrm = std::sqrt(nrm);
    const int osize = int(coeff.size());
    const double expected = 1.0 / std::sqrt(double(osize));
    for (auto& kv : coeff) {
      const cplx normalized = kv.second / (gauge * nrm);
      if (std::abs(std::abs(normalized) - expected) > 1.0e-10)
        throw OracleError(
            "incompatible symmetry request: symmetry maps do not define a one-dimensional orbit sector");
      b.lookup[kv.first] = Image{rep, normalized / expected, osize};
    }
    reps.push_back({rep, osize});
  }
  // parent rows are ascending states for every spin parent, and an
  // orbit-seed representative outside the parent sorts by its value: both
  // are the ascending order of the states
  std::sort(reps.begin(), reps.end());
  for (auto& r : reps) {
    b.states.push_back(r.first);
    b.orbit.push_back(r.second);
  }
}

// SpinBasisBuilder::build: src/basis/mod.rs:1312-1412 (spin one-half only)
static void build_basis(Basis& b) {
  if (b.L <= 0) throw OracleError("invalid Hilbert-space sector: symmetry sectors require at least one site");
  std::vector<u64> parents;
  if (!b.up_sectors.empty()) {
    for (int u : b.up_sectors) {
      auto part = fixed_weight_states(b.L, u);
      parents.insert(parents.end(), part.begin(), part.end());
    }
    std::sort(parents.begin(), parents.end());
  } else {
    parents = fixed_weight_states(b.L, b.up);
  }
  if (b.general()) {
    // build_spin_basis (bindings/capi/src/lib.rs:2849-2920)
    if (b.has_k || b.parity) throw OracleError("invalid options: built-in and general spin symmetries cannot be mixed");
    for (auto& g : b.gens) validate_map(g, b.L);
    general_sector(b, parents);
  } else {
    symmetry_sector(b, std::move(parents));
  }
  // ordered_basis: the reverse convention lists the rows in descending order
  // (bindings/capi/src/lib.rs:3412-3414)
  if (b.reverse) {
    std::reverse(b.states.begin(), b.states.end());
    std::reverse(b.orbit.begin(), b.orbit.end());
  }
  if (b.reverse || b.general())
    for (size_t i = 0; i < b.states.size(); ++i) b.rows[b.states[i]] = long(i);
}

// ------------------------------------------------------------- operator ----

struct Coupling {
  cplx coeff;
  std::vector<int> sites;
};
struct Term {
  std::string symbols;
  std::vector<Coupling> couplings;
};

struct Triplet {
  long row, col;
  cplx value;
};

struct Model {
  Basis basis;
  std::vector<Term> terms;
  std::vector<Triplet> entries;  // row-sorted
  bool real = true;
  long dim() const { return long(basis.states.size()); }
};

// Local action on one state, symbols applied right to left.
// visit_unreduced_local_transitions_with_symbols: src/basis/mod.rs:1122-1232
// (spin one-half: each symbol has at most one branch)
static bool local_action(const Basis& b, u64 state, const std::string& sym, const std::vector<int>& sites,
                         u64& out, cplx& amp) {
  double zf[2], raise[2], lower[2];
  double zscale = (b.norm == ANGULAR) ? 1.0 : 2.0;
  zf[0] = -0.5 * zscale;
  zf[1] = 0.5 * zscale;
  double ladder = (b.norm == PAULI) ? 2.0 : 1.0;
  raise[0] = 1.0 * ladder;
  raise[1] = 0.0;
  lower[0] = 0.0;
  lower[1] = 1.0 * ladder;
  double xyscale = (b.norm == PAULI_CARTESIAN) ? 1.0 : 0.5;
  amp = cplx(1.0, 0.0);
  u64 enc = state;
  for (int pos = int(sym.size()) - 1; pos >= 0; --pos) {
    int site = sites[pos];
    if (site < 0 || site >= b.L) {
      std::ostringstream os;
      os << "site " << site << " is outside a lattice with " << b.L << " sites";
      throw OracleError(os.str());
    }
    int digit = int((enc >> site) & 1);
    char op = sym[pos];
    switch (op) {
      case 'I': break;
      case 'z': amp *= zf[digit]; break;
      case '+':
        if (raise[digit] == 0.0) return false;
        enc += u64(1) << site;
        if (raise[digit] != 1.0) amp *= raise[digit];
        break;
      case '-':
        if (lower[digit] == 0.0) return false;
        enc -= u64(1) << site;
        if (lower[digit] != 1.0) amp *= lower[digit];
        break;
      case 'x':
      case 'y': {
        cplx rp = (op == 'x') ? cplx(xyscale, 0) : cplx(0, -xyscale);
        cplx lp = (op == 'x') ? cplx(xyscale, 0) : cplx(0, xyscale);
        if (raise[digit] != 0.0) {
          enc += u64(1) << site;
          amp *= rp * raise[digit];
        } else {
          enc -= u64(1) << site;
          amp *= lp * lower[digit];
        }
        break;
      }
      default: throw OracleError(std::string("invalid or unregistered operator: ") + op);
    }
  }
  out = enc;
  return true;
}

// operator_number_changes / selected_sectors_preserve_changes and the
// Cartesian decomposition of operator_sum_preserves_particle_sector:
// src/basis/mod.rs:294-371, src/operator/mod.rs:1963-2033
static bool preserves_sector(const Basis& b, const std::vector<Term>& terms) {
  struct Key {
    int degree;
    std::string op;
    std::vector<int> sites;
    bool operator<(const Key& o) const {
      if (degree != o.degree) return degree < o.degree;
      if (op != o.op) return op < o.op;
      return sites < o.sites;
    }
  };
  std::map<Key, cplx> comps;
  for (auto& t : terms) {
    for (auto& c : t.couplings) {
      std::vector<std::tuple<int, std::string, cplx>> parts{{0, "", cplx(1, 0)}};
      for (char s : t.symbols) {
        if (s == 'x' || s == 'y') {
          std::vector<std::tuple<int, std::string, cplx>> next;
          for (auto& [d, p, co] : parts) {
            cplx rp = (s == 'x') ? cplx(1, 0) : cplx(0, -1);
            cplx lp = (s == 'x') ? cplx(1, 0) : cplx(0, 1);
            next.push_back({d + 1, p + '+', co * rp});
            next.push_back({d + 1, p + '-', co * lp});
          }
          parts = next;
        } else if (s == 'I' || s == 'n' || s == 'z' || s == '+' || s == '-' || s == '|') {
          for (auto& pt : parts) std::get<1>(pt) += s;
        } else {
          throw OracleError("invalid or unregistered operator: " + t.symbols);
        }
      }
      for (auto& [d, p, co] : parts) comps[Key{d, p, c.sites}] += c.coeff * co;
    }
  }
  std::vector<int> sectors = b.up_sectors;
  bool fixed = b.up >= 0;
  if (!fixed && sectors.empty()) return true;
  for (auto& kv : comps) {
    if (std::abs(kv.second) <= EPS) continue;
    int change = 0;
    for (char s : kv.first.op) change += (s == '+') - (s == '-');
    auto contains = [&](int sec) {
      if (!sectors.empty()) return std::binary_search(sectors.begin(), sectors.end(), sec);
      return sec == b.up;
    };
    auto ok = [&](int src) {
      int tgt = src + change;
      return tgt < 0 || tgt > b.L || contains(tgt);
    };
    if (!sectors.empty()) {
      for (int s : sectors)
        if (!ok(s)) return false;
    } else if (!ok(b.up)) {
      return false;
    }
  }
  return true;
}

static cplx value_at(const std::vector<Triplet>& e, long row, long col) {
  auto it = std::lower_bound(e.begin(), e.end(), std::make_pair(row, col),
                             [](const Triplet& t, const std::pair<long, long>& k) {
                               return std::make_pair(t.row, t.col) < k;
                             });
  if (it != e.end() && it->row == row && it->col == col) return it->value;
  return cplx(0, 0);
}

// OperatorBuilder::build(MatrixFormat::MatrixFree): src/operator/mod.rs:2091-2216
// index_transition: src/basis/mod.rs:1561-1580
static void assemble(Model& m, bool check_herm, bool check_particles) {
  Basis& b = m.basis;
  if (check_particles && !preserves_sector(b, m.terms))
    throw OracleError("invalid Hilbert-space sector: operator sum does not preserve the selected particle sector");
  for (auto& t : m.terms)
    for (auto& c : t.couplings)
      if (c.sites.size() != t.symbols.size()) {
        std::ostringstream os;
        os << "invalid coupling: operator arity " << t.symbols.size() << " does not match " << c.sites.size()
           << " sites";
        throw OracleError(os.str());
      }
  m.entries.clear();
  std::vector<std::pair<long, cplx>> col;
  long n = m.dim();
  for (long column = 0; column < n; ++column) {
    col.clear();
    u64 src = b.states[column];
    int src_orbit = b.symmetric() ? b.orbit[column] : 1;
    for (auto& t : m.terms) {
      for (auto& c : t.couplings) {
        u64 tgt;
        cplx amp;
        if (!local_action(b, src, t.symbols, c.sites, tgt, amp)) continue;
        long row;
        cplx red(1, 0);
        if (b.symmetric()) {
          auto it = b.lookup.find(tgt);
          if (it == b.lookup.end()) continue;
          row = b.index(it->second.rep);
          red = std::sqrt(double(src_orbit) / double(it->second.orbit)) * std::conj(it->second.phase);
        } else {
          row = b.index(tgt);
          if (row < 0) continue;
        }
        col.push_back({row, c.coeff * amp * red});
      }
    }
    std::stable_sort(col.begin(), col.end(),
                     [](const std::pair<long, cplx>& a, const std::pair<long, cplx>& bb) { return a.first < bb.first; });
    size_t p = 0;
    while (p < col.size()) {
      long row = col[p].first;
      cplx v = col[p].second;
      ++p;
      while (p < col.size() && col[p].first == row) v += col[p++].second;
      if (std::abs(v) <= EPS) continue;
      m.entries.push_back({row, column, v});
    }
  }
  std::sort(m.entries.begin(), m.entries.end(), [](const Triplet& a, const Triplet& bb) {
    return std::make_pair(a.row, a.col) < std::make_pair(bb.row, bb.col);
  });
  m.real = true;
  for (auto& e : m.entries)
    if (e.value.imag() != 0.0) m.real = false;
  if (check_herm) {
    for (auto& e : m.entries)
      if (std::abs(e.value - std::conj(value_at(m.entries, e.col, e.row))) > 1.0e-12)
        throw OracleError("a Hermitian operator is required");
  }
}

// Operator::apply (MatrixFree COO loop): src/operator/mod.rs:1478-1482, 1541-1545
static void apply(const Model& m, const cplx* x, cplx* y) {
  long n = m.dim();
  for (long i = 0; i < n; ++i) y[i] = 0;
  if (m.real) {
    for (auto& e : m.entries) {
      y[e.row] += cplx(e.value.real() * x[e.col].real(), e.value.real() * x[e.col].imag());
    }
  } else {
    for (auto& e : m.entries) y[e.row] += e.value * x[e.col];
  }
}
static void apply_real(const Model& m, const double* x, double* y) {
  long n = m.dim();
  for (long i = 0; i < n; ++i) y[i] = 0;
  for (auto& e : m.entries) y[e.row] += e.value.real() * x[e.col];
}

// ------------------------------------------------------- dense eigen ------

// Cyclic Jacobi eigensolver for a complex Hermitian matrix (row-major, n×n).
// Eigenvalues ascending; eigenvectors as columns of vecs (row-major).
static void hermitian_eig(int n, std::vector<cplx> a, std::vector<double>& vals, std::vector<cplx>& vecs) {
  vecs.assign(size_t(n) * n, cplx(0, 0));
  for (int i = 0; i < n; ++i) vecs[size_t(i) * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double v = std::norm(a[size_t(i) * n + j]);
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * std::max(tot, 1e-300) || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        cplx apq = a[size_t(p) * n + q];
        double mag = std::abs(apq);
        if (mag < 1e-300) continue;
        double app = a[size_t(p) * n + p].real(), aqq = a[size_t(q) * n + q].real();
        cplx ph = apq / mag;  // a_pq = mag * ph
        double theta = 0.5 * std::atan2(2.0 * mag, aqq - app);
        double c = std::cos(theta), s = std::sin(theta);
        // Rotation in the (p,q) plane: columns p,q of G = [[c, s*ph],[-s*conj(ph), c]]
        cplx gpp = c, gpq = s * ph, gqp = -s * std::conj(ph), gqq = c;
        // A <- G^H A G
        for (int k = 0; k < n; ++k) {  // A G (columns)
          cplx akp = a[size_t(k) * n + p], akq = a[size_t(k) * n + q];
          a[size_t(k) * n + p] = akp * gpp + akq * gqp;
          a[size_t(k) * n + q] = akp * gpq + akq * gqq;
        }
        for (int k = 0; k < n; ++k) {  // G^H (rows)
          cplx apk = a[size_t(p) * n + k], aqk = a[size_t(q) * n + k];
          a[size_t(p) * n + k] = std::conj(gpp) * apk + std::conj(gqp) * aqk;
          a[size_t(q) * n + k] = std::conj(gpq) * apk + std::conj(gqq) * aqk;
        }
        a[size_t(p) * n + q] = 0;
        a[size_t(q) * n + p] = 0;
        for (int k = 0; k < n; ++k) {
          cplx vkp = vecs[size_t(k) * n + p], vkq = vecs[size_t(k) * n + q];
          vecs[size_t(k) * n + p] = vkp * gpp + vkq * gqp;
          vecs[size_t(k) * n + q] = vkp * gpq + vkq * gqq;
        }
      }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = a[size_t(i) * n + i].real();
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
  vals.resize(n);
  std::vector<cplx> sorted(size_t(n) * n);
  for (int j = 0; j < n; ++j) {
    vals[j] = d[order[j]];
    for (int i = 0; i < n; ++i) sorted[size_t(i) * n + j] = vecs[size_t(i) * n + order[j]];
  }
  vecs.swap(sorted);
}

// Real symmetric tridiagonal eigen-decomposition (via the Hermitian Jacobi
// routine); returns ascending values and column eigenvectors (row-major m×m).
static void tridiag_eig(const std::vector<double>& alpha, const std::vector<double>& beta, std::vector<double>& vals,
                        std::vector<double>& vecs) {
  int m = int(alpha.size());
  std::vector<cplx> a(size_t(m) * m, cplx(0, 0));
  for (int i = 0; i < m; ++i) {
    a[size_t(i) * m + i] = alpha[i];
    if (i + 1 < m) {
      a[size_t(i) * m + i + 1] = beta[i];
      a[size_t(i + 1) * m + i] = beta[i];
    }
  }
  std::vector<cplx> cv;
  hermitian_eig(m, a, vals, cv);
  vecs.resize(size_t(m) * m);
  // the eigenvectors of a real symmetric matrix from the complex rotations are
  // real up to a per-column phase; remove it
  for (int j = 0; j < m; ++j) {
    cplx ph(1, 0);
    double best = 0;
    for (int i = 0; i < m; ++i)
      if (std::abs(cv[size_t(i) * m + j]) > best) {
        best = std::abs(cv[size_t(i) * m + j]);
        ph = cv[size_t(i) * m + j] / best;
      }
    for (int i = 0; i < m; ++i) vecs[size_t(i) * m + j] = (cv[size_t(i) * m + j] / ph).real();
  }
}

// ---------------------------------------------------------------- eigsh ----

enum Target { SA = 0, LA = 1, SM = 2, LM = 3, BE = 4 };

struct EigshOptions {
  int k;
  int target;
  long krylov;  // <=0: none
  double tol = 1e-10;
  long max_iter = 1000;
  u64 seed = 0;
};

struct Eigensystem {
  std::vector<double> values;
  std::vector<std::vector<cplx>> vectors;
  std::vector<double> residuals;
  long iterations = 0;
  long passes = 0, second = 0;
  bool converged = true;
};

struct NonConvergence {
  long iterations;
  double residual;
};

// deterministic_start: src/solve/mod.rs:729-742
static std::vector<double> deterministic_start(long n, u64 seed) {
  u64 state = seed | 1;
  std::vector<double> v(n);
  for (long i = 0; i < n; ++i) {
    state = state * 6364136223846793005ULL + 1442695040888963407ULL;
    u64 mant = state >> 11;
    v[i] = double(mant) / double(u64(1) << 53) - 0.5;
  }
  double nrm = 0;
  for (double x : v) nrm += x * x;
  nrm = std::sqrt(nrm);
  for (double& x : v) x /= nrm;
  return v;
}

static int total_cmp(double a, double b) { return (a < b) ? -1 : (a > b) ? 1 : 0; }

// select_indices: src/solve/mod.rs:1002-1036
static std::vector<int> select_indices(const std::vector<double>& v, int target, int count) {
  int n = int(v.size());
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  if (target == BE) {
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return v[a] < v[b]; });
    int lower = count / 2, upper = count - lower;
    std::vector<int> sel(idx.begin(), idx.begin() + lower);
    for (int i = 0; i < upper; ++i) sel.push_back(idx[n - 1 - i]);
    std::stable_sort(sel.begin(), sel.end(), [&](int a, int b) { return v[a] < v[b]; });
    return sel;
  }
  auto key = [&](int i, int j) {
    int c = 0;
    switch (target) {
      case SA: c = total_cmp(v[i], v[j]); break;
      case LA: c = total_cmp(v[j], v[i]); break;
      case SM: c = total_cmp(std::abs(v[i]), std::abs(v[j])); break;
      case LM: c = total_cmp(std::abs(v[j]), std::abs(v[i])); break;
    }
    if (c != 0) return c < 0;
    return i < j;
  };
  std::sort(idx.begin(), idx.end(), key);
  idx.resize(std::min(count, n));
  return idx;
}

static bool target_less(int target, double a, double b) {
  switch (target) {
    case SA: return a < b;
    case LA: return b < a;
    case SM: return std::abs(a) < std::abs(b);
    case LM: return std::abs(b) < std::abs(a);
    default: return a < b;
  }
}

static cplx cinner(const std::vector<cplx>& a, const std::vector<cplx>& b) {
  cplx s(0, 0);
  for (size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
  return s;
}
static double cnorm(const std::vector<cplx>& a) {
  double s = 0;
  for (auto& x : a) s += std::norm(x);
  return std::sqrt(s);
}
static double rinner(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// lanczos_eigsh_real / lanczos_eigsh with DGKS: src/solve/mod.rs:1050-1099, 1116-1456
static Eigensystem lanczos_eigsh(const Model& m, const EigshOptions& o, const std::vector<cplx>* initial) {
  long n = m.dim();
  long req = o.krylov > 0 ? o.krylov : std::max<long>(8 * o.k + 64, 256);
  long kd = std::min(std::min(req, o.max_iter), n);
  if (kd <= o.k) throw OracleError("invalid options: the effective Krylov dimension must exceed eigenpairs");
  bool real_path = m.real;
  if (initial)
    for (auto& z : *initial)
      if (std::abs(z.imag()) > 1e-14) real_path = false;
  Eigensystem es;
  std::vector<std::vector<cplx>> basis;
  std::vector<double> alphas, betas;
  std::vector<cplx> first(n);
  if (initial) {
    first = *initial;
    if (real_path)
      for (auto& z : first) z = z.real();
    double nr = cnorm(first);
    if (!std::isfinite(nr) || nr <= EPS) throw NonConvergence{0, nr};
    for (auto& z : first) z /= nr;
  } else {
    auto d = deterministic_start(n, o.seed);
    for (long i = 0; i < n; ++i) first[i] = d[i];
  }
  basis.push_back(first);
  std::vector<cplx> out(n);
  const double DGKS = 0.7071067811865476;
  for (long it = 0; it < kd; ++it) {
    apply(m, basis[it].data(), out.data());
    double a = cinner(basis[it], out).real();
    alphas.push_back(a);
    for (long i = 0; i < n; ++i) out[i] -= a * basis[it][i];
    if (it > 0)
      for (long i = 0; i < n; ++i) out[i] -= betas[it - 1] * basis[it - 1][i];
    double before = cnorm(out);
    for (auto& v : basis) {
      cplx ov = cinner(v, out);
      for (long i = 0; i < n; ++i) out[i] -= ov * v[i];
    }
    double after = cnorm(out);
    bool second = false;
    if (before > EPS && after <= DGKS * before) {
      for (auto& v : basis) {
        cplx ov = cinner(v, out);
        for (long i = 0; i < n; ++i) out[i] -= ov * v[i];
      }
      after = cnorm(out);
      second = true;
    }
    es.passes += 1 + second;
    es.second += second;
    if (it + 1 == kd || after <= 1e-14) break;
    betas.push_back(after);
    for (auto& z : out) z /= after;
    basis.push_back(out);
  }
  long size = long(basis.size());
  if (size <= o.k) throw NonConvergence{size, INFINITY};
  std::vector<double> tv, tvec;
  std::vector<double> al(alphas.begin(), alphas.begin() + size), be(betas.begin(), betas.begin() + size - 1);
  tridiag_eig(al, be, tv, tvec);
  auto sel = select_indices(tv, o.target, o.k);
  struct Cand {
    double val;
    std::vector<cplx> vec;
    double res;
  };
  std::vector<Cand> cands;
  for (int idx : sel) {
    std::vector<cplx> v(n, cplx(0, 0));
    for (long j = 0; j < size; ++j) {
      double c = tvec[size_t(j) * size + idx];
      for (long i = 0; i < n; ++i) v[i] += c * basis[j][i];
    }
    double nr = cnorm(v);
    for (auto& z : v) z /= nr;
    apply(m, v.data(), out.data());
    double ev = cinner(v, out).real();
    double r = 0;
    for (long i = 0; i < n; ++i) r += std::norm(out[i] - ev * v[i]);
    cands.push_back({ev, v, std::sqrt(r)});
  }
  std::stable_sort(cands.begin(), cands.end(), [&](const Cand& a, const Cand& b) {
    return target_less(o.target == BE ? SA : o.target, a.val, b.val);
  });
  double fail = 0;
  for (auto& c : cands)
    if (c.res > o.tol * std::max(std::abs(c.val), 1.0)) fail = std::max(fail, c.res);
  if (fail > 0) throw NonConvergence{size, fail};
  for (auto& c : cands) {
    es.values.push_back(c.val);
    es.vectors.push_back(c.vec);
    es.residuals.push_back(c.res);
  }
  es.iterations = size;
  return es;
}

// dense_eigsh: src/solve/mod.rs:2198-2240 (materialize, Hermiticity, eigen, RQ refine)
static Eigensystem dense_eigsh(const Model& m, const EigshOptions& o) {
  long n = m.dim();
  std::vector<cplx> a(size_t(n) * n, cplx(0, 0));
  for (auto& e : m.entries) a[size_t(e.row) * n + e.col] = e.value;
  for (long i = 0; i < n; ++i)
    for (long j = 0; j < n; ++j)
      if (std::abs(a[size_t(i) * n + j] - std::conj(a[size_t(j) * n + i])) > 1e-12)
        throw OracleError("a Hermitian operator is required");
  std::vector<double> vals;
  std::vector<cplx> vecs;
  hermitian_eig(int(n), a, vals, vecs);
  auto sel = select_indices(vals, o.target, o.k);
  Eigensystem es;
  std::vector<cplx> out(n);
  for (int idx : sel) {
    std::vector<cplx> v(n);
    for (long i = 0; i < n; ++i) v[i] = vecs[size_t(i) * n + idx];
    apply(m, v.data(), out.data());
    double ns = cinner(v, v).real();
    double ev = cinner(v, out).real() / ns;
    double r = 0;
    for (long i = 0; i < n; ++i) r += std::norm(out[i] - ev * v[i]);
    es.values.push_back(ev);
    es.vectors.push_back(v);
    es.residuals.push_back(std::sqrt(r));
  }
  es.iterations = 1;
  return es;
}

// ------------------------------------------------------- thick restart ----

// orthonormalize_{real,complex}_with_projection: src/solve/mod.rs:1519-1586
// (complex :1588-1641). Modified Gram-Schmidt against the locked vectors,
// then the basis (recording the basis overlaps of the first pass), a DGKS
// second pass when the norm fell to 1/sqrt(2) of its start, and a drop when
// the remainder is not finite or at most 1e-14.
static bool orthonormalize(const std::vector<std::vector<cplx>>& locked, const std::vector<std::vector<cplx>>& basis,
                           std::vector<cplx>& v, Eigensystem& es, std::vector<cplx>* projection) {
  const double DGKS = 0.7071067811865476;
  const size_t n = v.size();
  const double before = cnorm(v);
  auto pass = [&](bool record) {
    for (auto& l : locked) {
      const cplx ov = cinner(l, v);
      for (size_t i = 0; i < n; ++i) v[i] -= ov * l[i];
    }
    for (auto& b : basis) {
      const cplx ov = cinner(b, v);
      if (record && projection) projection->push_back(ov);
      for (size_t i = 0; i < n; ++i) v[i] -= ov * b[i];
    }
  };
  if (projection) projection->clear();
  pass(true);
  double nr = cnorm(v);
  const bool second = before > EPS && nr <= DGKS * before;
  if (second) {
    pass(false);
    nr = cnorm(v);
  }
  es.passes += 1 + (second ? 1 : 0);
  es.second += second ? 1 : 0;
  if (!std::isfinite(nr) || nr <= 1e-14) return false;
  for (auto& z : v) z /= nr;
  return true;
}

// thick_restarted_eigsh_{real,complex}: src/solve/mod.rs:1651-1913 (complex
// :1914-2167), window thick_restart_dimension :1644-1649. The real route is
// taken for real operators and real initial vectors (:2169-2196); here both
// run in complex arithmetic, the real one on vectors with zero imaginary
// parts and Ritz vectors made real.
static Eigensystem thick_restart(const Model& m, const EigshOptions& o, const std::vector<std::vector<cplx>>& init) {
  const long n = m.dim();
  const long rd = std::min(o.krylov > 0 ? o.krylov : std::max<long>(4L * o.k + 24, 32), n);
  if (rd <= o.k) throw OracleError("invalid options: the effective Krylov dimension must exceed eigenpairs");
  bool real = m.real;
  for (auto& v : init)
    for (auto& z : v)
      if (std::abs(z.imag()) > 1e-14) real = false;
  struct Kept {
    std::vector<cplx> vec, act;
    bool has_act;
  };
  struct Cand {
    double value;
    std::vector<cplx> vec;
    double residual;
    std::vector<cplx> act;
  };
  std::vector<Kept> retained;
  for (auto& v : init) {
    Kept k{v, {}, false};
    if (real)
      for (auto& z : k.vec) z = z.real();
    retained.push_back(std::move(k));
  }
  std::vector<Cand> locked;
  Eigensystem es;
  long iterations = 0;
  double last_residual = INFINITY;
  u64 cycle = 0;
  auto locked_vectors = [&]() {
    std::vector<std::vector<cplx>> lv;
    for (auto& c : locked) lv.push_back(c.vec);
    return lv;
  };
  while (iterations < o.max_iter && long(locked.size()) < o.k) {
    const auto lv = locked_vectors();
    std::vector<std::vector<cplx>> basis, applied;
    bool reuse = !retained.empty();
    for (auto& r : retained) reuse = reuse && r.has_act;
    for (auto& r : retained) {
      if (reuse) {
        const double nr = cnorm(r.vec);
        if (std::isfinite(nr) && nr > 1e-14) {
          for (auto& z : r.vec) z /= nr;
          for (auto& z : r.act) z /= nr;
          basis.push_back(r.vec);
          applied.push_back(r.act);
        }
      } else {
        std::vector<cplx> v = r.vec;
        if (orthonormalize(lv, basis, v, es, nullptr)) basis.push_back(v);
      }
      if (long(basis.size()) >= rd / 2) break;
    }
    retained.clear();
    if (basis.empty()) {
      auto d = deterministic_start(n, o.seed + cycle * 0x9e3779b9ULL);
      std::vector<cplx> v(n);
      for (long i = 0; i < n; ++i) v[i] = d[i];
      if (orthonormalize(lv, basis, v, es, nullptr)) basis.push_back(v);
    }
    if (basis.empty()) break;
    const size_t known = applied.size();
    std::vector<std::vector<cplx>> columns;
    std::vector<cplx> proj;
    if (reuse) {
      for (size_t c = 0; c < known; ++c) {
        if (long(basis.size()) >= rd) break;
        std::vector<cplx> v = applied[c];
        const bool keep = orthonormalize(lv, basis, v, es, &proj);
        columns.emplace_back(proj.begin(), proj.begin() + long(c) + 1);
        if (keep) basis.push_back(v);
      }
    }
    size_t cursor = known;
    std::vector<cplx> out(n);
    while (cursor < basis.size() && iterations < o.max_iter && long(applied.size()) < rd) {
      apply(m, basis[cursor].data(), out.data());
      ++iterations;
      applied.push_back(out);
      std::vector<cplx> v = out;
      const bool keep = orthonormalize(lv, basis, v, es, &proj);
      columns.emplace_back(proj.begin(), proj.begin() + long(cursor) + 1);
      if (long(basis.size()) < rd && keep) basis.push_back(v);
      ++cursor;
    }
    basis.resize(applied.size());
    if (basis.empty()) break;
    const int size = int(basis.size());
    std::vector<cplx> P(size_t(size) * size, cplx(0, 0));
    for (int c = 0; c < int(columns.size()) && c < size; ++c)
      for (int r = 0; r <= c && r < int(columns[size_t(c)].size()); ++r) {
        const cplx v = real ? cplx(columns[size_t(c)][size_t(r)].real(), 0) : columns[size_t(c)][size_t(r)];
        P[size_t(r) * size + c] = v;
        P[size_t(c) * size + r] = std::conj(v);
      }
    std::vector<double> vals;
    std::vector<cplx> U;
    hermitian_eig(size, P, vals, U);
    const long active = o.k - long(locked.size());
    const long keep_limit = std::max<long>(std::min<long>(active, rd / 2), 1);
    const auto idx = select_indices(vals, o.target, int(std::min<long>(keep_limit, size)));
    std::vector<Cand> cands;
    for (int q : idx) {
      cplx ph(1, 0);
      if (real) {  // the eigenvector of a real symmetric matrix, made real
        double best = 0;
        for (int j = 0; j < size; ++j)
          if (std::abs(U[size_t(j) * size + q]) > best) {
            best = std::abs(U[size_t(j) * size + q]);
            ph = U[size_t(j) * size + q] / best;
          }
      }
      std::vector<cplx> vec(n, cplx(0, 0)), act(n, cplx(0, 0));
      for (int j = 0; j < size; ++j) {
        cplx c = U[size_t(j) * size + q] / ph;
        if (real) c = c.real();
        for (long i = 0; i < n; ++i) {
          vec[i] += c * basis[size_t(j)][i];
          act[i] += c * applied[size_t(j)][i];
        }
      }
      const double rn = cnorm(vec);
      if (!std::isfinite(rn) || rn <= EPS) throw NonConvergence{iterations, rn};
      for (auto& z : vec) z /= rn;
      for (auto& z : act) z /= rn;
      const double value = cinner(vec, act).real();
      double r2 = 0;
      for (long i = 0; i < n; ++i) r2 += std::norm(act[i] - value * vec[i]);
      cands.push_back({value, vec, std::sqrt(r2), act});
    }
    // sort_real_candidates (stable): src/solve/mod.rs:1484-1508
    std::stable_sort(cands.begin(), cands.end(), [&](const Cand& a, const Cand& b) {
      return target_less(o.target == BE ? SA : o.target, a.value, b.value);
    });
    last_residual = 0;
    for (long r = 0; r < active && r < long(cands.size()); ++r) last_residual = std::max(last_residual, cands[size_t(r)].residual);
    std::vector<Kept> next;
    for (size_t r = 0; r < cands.size(); ++r) {
      auto& c = cands[r];
      const bool conv = c.residual <= o.tol * std::max(std::abs(c.value), 1.0);
      if (long(r) < active && conv && long(locked.size()) < o.k) {
        std::vector<cplx> v = c.vec;
        if (orthonormalize(locked_vectors(), {}, v, es, nullptr)) locked.push_back({c.value, v, c.residual, {}});
      } else if (long(next.size()) < keep_limit) {
        next.push_back({c.vec, c.act, true});
      }
    }
    retained = std::move(next);
    ++cycle;
  }
  if (long(locked.size()) < o.k) throw NonConvergence{iterations, last_residual};
  std::stable_sort(locked.begin(), locked.end(), [&](const Cand& a, const Cand& b) {
    return target_less(o.target == BE ? SA : o.target, a.value, b.value);
  });
  locked.resize(size_t(o.k));
  for (auto& c : locked) {
    es.values.push_back(c.value);
    es.vectors.push_back(c.vec);
    es.residuals.push_back(c.residual);
  }
  es.iterations = iterations;
  return es;
}

// the thick restart and its dense fallback for small full windows:
// src/solve/mod.rs:2306-2327
static Eigensystem thick_restart_route(const Model& m, const EigshOptions& o,
                                       const std::vector<std::vector<cplx>>& init) {
  const long n = m.dim();
  try {
    return thick_restart(m, o, init);
  } catch (NonConvergence& nc) {
    if (n <= 2048 && o.max_iter >= n && (o.krylov <= 0 || o.krylov == n)) {
      auto r = dense_eigsh(m, o);
      r.iterations += nc.iterations;
      return r;
    }
    throw;
  }
}

// eigsh + expanding_lanczos_eigsh dispatch: src/solve/mod.rs:2242-2350, 364-372, 425-448
static Eigensystem eigsh(const Model& m, EigshOptions o, const std::vector<cplx>* initial) {
  long n = m.dim();
  if (o.k <= 0 || o.k >= n)
    throw OracleError("invalid options: eigenpairs must be positive and smaller than the operator dimension");
  if (!std::isfinite(o.tol) || o.tol <= 0 || o.max_iter <= 0)
    throw OracleError("invalid options: tolerance and max_iterations must be positive");
  if (o.krylov > 0 && (o.krylov <= o.k || o.krylov > n))
    throw OracleError("invalid options: krylov_dimension must exceed eigenpairs and not exceed dimension");
  bool dense = n <= 128 || (n <= 256 && o.krylov <= 0);
  if (dense) return dense_eigsh(m, o);
  if (o.krylov <= 0) {
    long sub = std::min(std::min(std::max<long>(8 * o.k + 64, 256), n), o.max_iter);
    long spent = 0;
    for (;;) {
      EigshOptions att = o;
      att.krylov = sub;
      att.max_iter = sub;
      try {
        auto r = lanczos_eigsh(m, att, initial);
        r.iterations += spent;
        return r;
      } catch (NonConvergence& nc) {
        spent += nc.iterations;
        if (sub == n && n <= 2048) {
          auto r = dense_eigsh(m, o);
          r.iterations += spent;
          return r;
        }
        long remaining = o.max_iter > spent ? o.max_iter - spent : 0;
        long next = std::min(std::min(sub * 2, n), remaining);
        if (next <= sub) throw nc;
        sub = next;
      }
    }
  }
  if (o.krylov >= 96) return lanczos_eigsh(m, o, initial);
  std::vector<std::vector<cplx>> sub;
  if (initial) sub.push_back(*initial);
  return thick_restart_route(m, o, sub);
}

// eigsh_with_initial_subspace: src/solve/mod.rs:2385-2410 (validation done by
// the caller); one or no vector runs the single-start dispatch above
static Eigensystem eigsh_subspace(const Model& m, EigshOptions o, const std::vector<std::vector<cplx>>& sub) {
  long n = m.dim();
  if (sub.size() <= 1) return eigsh(m, o, sub.empty() ? nullptr : &sub[0]);
  if (o.k <= 0 || o.k >= n)
    throw OracleError("invalid options: eigenpairs must be positive and smaller than the operator dimension");
  if (!std::isfinite(o.tol) || o.tol <= 0 || o.max_iter <= 0)
    throw OracleError("invalid options: tolerance and max_iterations must be positive");
  if (o.krylov > 0 && (o.krylov <= o.k || o.krylov > n))
    throw OracleError("invalid options: krylov_dimension must exceed eigenpairs and not exceed dimension");
  if (n <= 128 || (n <= 256 && o.krylov <= 0)) return dense_eigsh(m, o);
  // expanding_lanczos_eigsh with a genuine multi-vector start: the
  // unrestarted window takes only one start, so every multi-vector request
  // goes to the thick restart (src/solve/mod.rs:2249-2304)
  return thick_restart_route(m, o, sub);
}

// balanced_subspace_start: src/solve/mod.rs:708-727
static std::vector<cplx> balanced_subspace_start(const std::vector<std::vector<cplx>>& sub) {
  if (sub.empty()) throw OracleError("invalid options: initial subspace must not be empty");
  const size_t n = sub[0].size();
  std::vector<cplx> out(n, cplx(0, 0));
  const double count = double(sub.size());
  for (size_t j = 0; j < sub.size(); ++j) {
    const cplx c = std::polar(1.0 / std::sqrt(count), 2.0 * M_PI * double(j) / count);
    for (size_t i = 0; i < n; ++i) out[i] += c * sub[j][i];
  }
  double nr = cnorm(out);
  if (!std::isfinite(nr) || nr <= EPS) {
    out = sub[0];
    nr = cnorm(out);
    if (!std::isfinite(nr) || nr <= EPS) throw NonConvergence{0, nr};
  }
  for (auto& z : out) z /= nr;
  return out;
}

// eigsh_with_workspace: src/solve/mod.rs:2419-2444 (the workspace holds the
// previous solve's eigenvectors; its dimension matches the operator's)
static Eigensystem eigsh_workspace(const Model& m, EigshOptions o, const std::vector<std::vector<cplx>>& ws) {
  if (ws.empty()) return eigsh(m, o, nullptr);
  if ((o.krylov <= 0 || o.krylov >= 96) && ws.size() > 1) {
    auto start = balanced_subspace_start(ws);
    return eigsh(m, o, &start);
  }
  return eigsh_subspace(m, o, ws);
}

// --------------------------------------------------------- evolution ------

struct Projection {
  double init_norm = 0;
  std::vector<std::vector<cplx>> basis;
  double residual_beta = 0;
  std::vector<double> evals;
  std::vector<double> evecs;  // row-major size×size
};

// lanczos_projection(_real): src/solve/mod.rs:2909-3054
static Projection lanczos_projection(const Model& m, const std::vector<cplx>& init, long dimension) {
  Projection p;
  long n = long(init.size());
  bool realp = m.real;
  for (auto& z : init)
    if (z.imag() != 0.0) realp = false;
  if (realp) {
    double s = 0;
    for (auto& z : init) s += z.real() * z.real();
    p.init_norm = std::sqrt(s);
  } else {
    p.init_norm = cnorm(init);
    if (p.init_norm <= EPS) return p;
  }
  long kd = std::max<long>(std::min(dimension, n), 1);
  std::vector<cplx> first(n);
  for (long i = 0; i < n; ++i) first[i] = realp ? cplx(init[i].real() / p.init_norm, 0) : init[i] / p.init_norm;
  p.basis.push_back(first);
  std::vector<double> alphas, betas;
  std::vector<cplx> out(n);
  for (long it = 0; it < kd; ++it) {
    apply(m, p.basis[it].data(), out.data());
    cplx al = cinner(p.basis[it], out);
    if (!realp && std::abs(al.imag()) > 1e-10) throw OracleError("a Hermitian operator is required");
    alphas.push_back(al.real());
    for (long i = 0; i < n; ++i) out[i] -= al.real() * p.basis[it][i];
    if (it > 0)
      for (long i = 0; i < n; ++i) out[i] -= betas[it - 1] * p.basis[it - 1][i];
    double beta = cnorm(out);
    p.residual_beta = beta;
    if (it + 1 == kd || beta <= 1e-14) break;
    betas.push_back(beta);
    for (auto& z : out) z /= beta;
    p.basis.push_back(out);
  }
  long size = long(p.basis.size());
  std::vector<double> be(betas.begin(), betas.begin() + (size - 1));
  tridiag_eig(alphas, be, p.evals, p.evecs);
  return p;
}

// projected_exponential_coefficients/_action: src/solve/mod.rs:3056-3129
static std::vector<cplx> projected_action(const Projection& p, double interval, long n) {
  long size = long(p.basis.size());
  std::vector<cplx> out(n, cplx(0, 0));
  if (size == 0) return out;
  std::vector<cplx> coef(size, cplx(0, 0));
  for (long e = 0; e < size; ++e) {
    cplx ex = std::exp(cplx(0.0, -interval * p.evals[e]));
    cplx w = p.init_norm * p.evecs[e] * ex;  // evecs[(0, e)]
    for (long b = 0; b < size; ++b) coef[b] += p.evecs[size_t(b) * size + e] * w;
  }
  for (long b = 0; b < size; ++b)
    for (long i = 0; i < n; ++i) out[i] += coef[b] * p.basis[b][i];
  double on = cnorm(out);
  if (on > EPS && std::isfinite(on)) {
    double sc = p.init_norm / on;
    for (auto& z : out) z *= sc;
  }
  return out;
}

// projected_residual_amplitude + projected_exponential_error_bound: src/solve/mod.rs:3131-3195
static double residual_amplitude(const Projection& p, double interval) {
  long size = long(p.basis.size());
  if (size == 0) return 0;
  cplx a(0, 0);
  for (long e = 0; e < size; ++e)
    a += p.init_norm * p.evecs[e] * p.evecs[size_t(size - 1) * size + e] * std::exp(cplx(0, -interval * p.evals[e]));
  return std::abs(a);
}
static double error_bound(const Projection& p, double interval, long n) {
  double dur = std::abs(interval);
  if (dur <= EPS || p.basis.empty() || long(p.basis.size()) == n || p.residual_beta <= 1e-14) return 0;
  double mn = INFINITY, mx = -INFINITY;
  for (double v : p.evals) {
    mn = std::min(mn, v);
    mx = std::max(mx, v);
  }
  double span = dur * std::max(mx - mn, 0.0);
  long osc = long(std::min(std::ceil(span / M_PI), 60.0));
  long intervals = std::min<long>(16 + 4 * osc, 256);
  if (intervals % 2) ++intervals;
  double dir = interval > 0 ? 1.0 : (interval < 0 ? -1.0 : 0.0);
  double step = dur / double(intervals);
  double integral = 0;
  for (long i = 0; i <= intervals; ++i) {
    double w = (i == 0 || i == intervals) ? 1.0 : (i % 2 == 0 ? 2.0 : 4.0);
    integral += w * residual_amplitude(p, dir * step * double(i));
  }
  return 1.25 * p.residual_beta * step * integral / 3.0;
}

struct EvolveStats {
  long projections = 0, matvecs = 0, accepted = 0, rejected = 0;
  double max_error = 0;
};

// evolve_hermitian_adaptive_grid: src/solve/mod.rs:3197-3316
static std::vector<std::vector<cplx>> evolve(const Model& m, const std::vector<cplx>& init,
                                             const std::vector<double>& times, long krylov, double tol,
                                             long max_substeps, EvolveStats& st) {
  if (times.empty()) throw OracleError("invalid options: times must be a nonempty finite nondecreasing grid");
  for (size_t i = 0; i < times.size(); ++i)
    if (!std::isfinite(times[i]) || (i > 0 && times[i - 1] > times[i]))
      throw OracleError("invalid options: times must be a nonempty finite nondecreasing grid");
  if (krylov <= 0 || !std::isfinite(tol) || tol <= 0 || max_substeps <= 0)
    throw OracleError("invalid options: evolution controls must be positive");
  long n = long(init.size());
  std::vector<std::vector<cplx>> states;
  std::vector<cplx> cur = init;
  double t = 0;
  size_t oi = 0;
  long substeps = 0;
  double local_tol = tol / double(std::max<size_t>(times.size(), 1));
  while (oi < times.size()) {
    double target = times[oi];
    if (std::abs(target - t) <= 16 * EPS * std::max(std::abs(target), 1.0)) {
      states.push_back(cur);
      t = target;
      ++oi;
      continue;
    }
    if (substeps >= max_substeps) {
      std::ostringstream os;
      os << "solver did not converge after " << substeps << " iterations";
      throw OracleError(os.str());
    }
    Projection p = lanczos_projection(m, cur, krylov);
    st.projections++;
    st.matvecs += long(p.basis.size());
    double scale = std::max(cnorm(cur), 1.0);
    double thr = local_tol * scale;
    std::vector<std::pair<double, std::vector<cplx>>> accepted;
    for (size_t j = oi; j < times.size(); ++j) {
      double iv = times[j] - t;
      double err = error_bound(p, iv, n);
      st.max_error = std::max(st.max_error, err);
      if (err > thr) {
        st.rejected++;
        break;
      }
      accepted.push_back({times[j], projected_action(p, iv, n)});
    }
    if (!accepted.empty()) {
      for (auto& a : accepted) states.push_back(a.second);
      t = accepted.back().first;
      cur = accepted.back().second;
      oi += accepted.size();
      substeps++;
      st.accepted++;
      continue;
    }
    double tiv = target - t;
    double dir = tiv > 0 ? 1.0 : -1.0;
    double min_iv = 16 * EPS * std::max(std::abs(target), 1.0);
    double rej = std::abs(tiv), acc = rej;
    for (;;) {
      double err = error_bound(p, dir * acc, n);
      st.max_error = std::max(st.max_error, err);
      if (err <= thr || acc <= min_iv) break;
      st.rejected++;
      rej = acc;
      acc *= 0.5;
    }
    if (acc > min_iv && rej > acc) {
      for (int r = 0; r < 8; ++r) {
        double trial = 0.5 * (acc + rej);
        double err = error_bound(p, dir * trial, n);
        st.max_error = std::max(st.max_error, err);
        if (err <= thr) acc = trial;
        else {
          st.rejected++;
          rej = trial;
        }
      }
    }
    double iv = dir * acc;
    t += iv;
    cur = projected_action(p, iv, n);
    substeps++;
    st.accepted++;
  }
  return states;
}

// ------------------------------------------------------- Taylor plan ------

// Al-Mohy & Higham (2011), Table 3.1 theta_m for u = 2^-53, as tabulated at
// src/solve/mod.rs:14-50.
static const std::pair<int, double> THETA[] = {
    {1, 2.29e-16}, {2, 2.58e-8}, {3, 1.39e-5}, {4, 3.40e-4}, {5, 2.40e-3}, {6, 9.07e-3}, {7, 2.38e-2},
    {8, 5.00e-2},  {9, 8.96e-2}, {10, 1.44e-1}, {11, 2.14e-1}, {12, 3.00e-1}, {13, 4.00e-1}, {14, 5.14e-1},
    {15, 6.41e-1}, {16, 7.81e-1}, {17, 9.31e-1}, {18, 1.09}, {19, 1.26}, {20, 1.44}, {21, 1.62},
    {22, 1.82},    {23, 2.01},   {24, 2.22},   {25, 2.43},   {26, 2.64},   {27, 2.86},   {28, 3.08},
    {29, 3.31},    {30, 3.54},   {35, 4.7},    {40, 6.0},    {45, 7.2},    {50, 8.5},    {55, 9.9}};

struct Plan {
  long n;
  cplx coeff, shift;
  double one_norm;
  int degree;
  long scaling;
  double tol;
};

// shifted_trace_and_one_norm + ExpmActionPlan::new: src/solve/mod.rs:59-169
static Plan make_plan(const Model& m, cplx coeff, int max_degree, double tol, long max_substeps) {
  if (!std::isfinite(coeff.real()) || !std::isfinite(coeff.imag()) || max_degree == 0 || !std::isfinite(tol) ||
      tol <= 0 || max_substeps == 0)
    throw OracleError("invalid options: invalid exponential coefficient or numerical controls");
  long n = m.dim();
  std::vector<cplx> diag(n, cplx(0, 0));
  std::vector<double> colsum(n, 0.0);
  for (auto& e : m.entries) {
    if (e.row == e.col) diag[e.row] += e.value;
    else colsum[e.col] += std::abs(e.value);
  }
  cplx trace(0, 0);
  for (auto& d : diag) trace += d;
  cplx shift = n ? trace / double(n) : cplx(0, 0);
  for (long c = 0; c < n; ++c) colsum[c] += std::abs(diag[c] - shift);
  double nrm = 0;
  for (double v : colsum) nrm = std::max(nrm, v);
  double scaled = std::abs(coeff) * nrm;
  int degree = 0;
  long scaling = 1;
  if (scaled != 0.0) {
    bool found = false;
    long best = 0;
    for (auto& th : THETA) {
      if (th.first > max_degree) continue;
      long s = long(std::max(std::ceil(scaled / th.second), 1.0));
      long cost = long(th.first) * s;
      if (!found || cost < best) {
        found = true;
        best = cost;
        degree = th.first;
        scaling = s;
      }
    }
    if (!found) throw OracleError("invalid options: Taylor degree is below the minimum supported degree");
  }
  if (scaling > max_substeps) {
    std::ostringstream os;
    os << "solver did not converge after " << max_substeps << " iterations";
    throw OracleError(os.str());
  }
  return Plan{n, coeff, shift, nrm, degree, scaling, tol};
}

static double inf_norm(const std::vector<cplx>& v) {
  double m = 0;
  for (auto& z : v) m = std::max(m, std::abs(z));
  return m;
}

// ExpmActionPlan::apply: src/solve/mod.rs:198-247
static std::vector<cplx> plan_apply(const Model& m, const Plan& p, const std::vector<cplx>& init) {
  if (std::abs(p.coeff) <= EPS || inf_norm(init) == 0.0) return init;
  cplx factor = p.coeff / double(p.scaling);
  cplx eta = std::exp(factor * p.shift);
  std::vector<cplx> state = init, applied(p.n);
  for (long s = 0; s < p.scaling; ++s) {
    std::vector<cplx> term = state, sum = state;
    double prev = inf_norm(term);
    for (int order = 1; order <= p.degree; ++order) {
      apply(m, term.data(), applied.data());
      cplx sc = factor / double(order);
      for (long i = 0; i < p.n; ++i) applied[i] = sc * (applied[i] - p.shift * term[i]);
      std::swap(term, applied);
      for (long i = 0; i < p.n; ++i) sum[i] += term[i];
      double tn = inf_norm(term);
      if (prev + tn <= p.tol * std::max(inf_norm(sum), std::numeric_limits<double>::min())) break;
      prev = tn;
    }
    for (long i = 0; i < p.n; ++i) state[i] = eta * sum[i];
  }
  return state;
}

}  // namespace oracle

// ------------------------------------------------------------ C entry ------
// Flat C interface for the test harness (ctypes). Every function returns 0 on
// success and a negative value on error; the last error text is available
// from oracle_last_error().

using namespace oracle;

static thread_local std::string g_err;

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// terms_text: one coupling per line: "<symbols> <re> <im> <site> <site> ..."
// Couplings that share a symbol string and appear consecutively form one term.
void* oracle_model_new(int sites, int up, const int* up_sectors, int n_up_sectors, int has_k, int k, int parity,
                       int norm, const char* terms_text, int check_herm, int check_particles) {
  try {
    auto* m = new Model();
    m->basis.L = sites;
    m->basis.up = up;
    for (int i = 0; i < n_up_sectors; ++i) m->basis.up_sectors.push_back(up_sectors[i]);
    std::sort(m->basis.up_sectors.begin(), m->basis.up_sectors.end());
    m->basis.up_sectors.erase(std::unique(m->basis.up_sectors.begin(), m->basis.up_sectors.end()),
                              m->basis.up_sectors.end());
    m->basis.has_k = has_k != 0;
    m->basis.k = k;
    m->basis.parity = parity;
    m->basis.norm = norm;
    build_basis(m->basis);
    std::istringstream in(terms_text ? terms_text : "");
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string sym;
      double re, im;
      if (!(ls >> sym >> re >> im)) continue;
      std::vector<int> s;
      int x;
      while (ls >> x) s.push_back(x);
      if (m->terms.empty() || m->terms.back().symbols != sym) m->terms.push_back(Term{sym, {}});
      m->terms.back().couplings.push_back(Coupling{cplx(re, im), s});
    }
    if (!m->terms.empty()) assemble(*m, check_herm != 0, check_particles != 0);
    return m;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

// The general symmetry sector: nsym generators, dest [nsym][sites], local
// [nsym][sites][2], sectors [nsym]; reverse: maximum representatives and
// descending rows (with nsym = 0: the rows of the plain or built-in sector
// reversed).
void* oracle_model_new_general(int sites, int up, const int* up_sectors, int n_up_sectors, int has_k, int k,
                               int parity, int norm, const char* terms_text, int check_herm, int check_particles,
                               int nsym, const int* dest, const int* local, const int* sectors, int reverse) {
  try {
    auto* m = new Model();
    m->basis.L = sites;
    m->basis.up = up;
    for (int i = 0; i < n_up_sectors; ++i) m->basis.up_sectors.push_back(up_sectors[i]);
    std::sort(m->basis.up_sectors.begin(), m->basis.up_sectors.end());
    m->basis.up_sectors.erase(std::unique(m->basis.up_sectors.begin(), m->basis.up_sectors.end()),
                              m->basis.up_sectors.end());
    m->basis.has_k = has_k != 0;
    m->basis.k = k;
    m->basis.parity = parity;
    m->basis.norm = norm;
    m->basis.reverse = reverse != 0;
    for (int g = 0; g < nsym; ++g) {
      GeneralMap gm;
      for (int i = 0; i < sites; ++i) {
        gm.dest.push_back(dest[size_t(g) * size_t(sites) + size_t(i)]);
        gm.local.push_back({local[(size_t(g) * size_t(sites) + size_t(i)) * 2],
                            local[(size_t(g) * size_t(sites) + size_t(i)) * 2 + 1]});
      }
      gm.sector = sectors[g];
      m->basis.gens.push_back(gm);
    }
    build_basis(m->basis);
    std::istringstream in(terms_text ? terms_text : "");
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string sym;
      double re, im;
      if (!(ls >> sym >> re >> im)) continue;
      std::vector<int> s;
      int x;
      while (ls >> x) s.push_back(x);
      if (m->terms.empty() || m->terms.back().symbols != sym) m->terms.push_back(Term{sym, {}});
      m->terms.back().couplings.push_back(Coupling{cplx(re, im), s});
    }
    if (!m->terms.empty()) assemble(*m, check_herm != 0, check_particles != 0);
    return m;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void oracle_model_free(void* h) { delete static_cast<Model*>(h); }

long oracle_dim(void* h) { return static_cast<Model*>(h)->dim(); }
long oracle_nnz(void* h) { return long(static_cast<Model*>(h)->entries.size()); }
int oracle_is_real(void* h) { return static_cast<Model*>(h)->real ? 1 : 0; }

void oracle_states(void* h, unsigned long long* states, int* orbits) {
  auto* m = static_cast<Model*>(h);
  for (size_t i = 0; i < m->basis.states.size(); ++i) {
    states[i] = m->basis.states[i];
    if (orbits) orbits[i] = m->basis.orbit[i];
  }
}

// reduction image of one parent state: returns 1 if present, 0 if not
int oracle_image(void* h, unsigned long long state, unsigned long long* rep, double* phase_re, double* phase_im,
                 int* orbit) {
  auto* m = static_cast<Model*>(h);
  if (!m->basis.symmetric()) {
    if (m->basis.index(state) < 0) return 0;
    *rep = state;
    *phase_re = 1;
    *phase_im = 0;
    *orbit = 1;
    return 1;
  }
  auto it = m->basis.lookup.find(state);
  if (it == m->basis.lookup.end()) return 0;
  *rep = it->second.rep;
  *phase_re = it->second.phase.real();
  *phase_im = it->second.phase.imag();
  *orbit = it->second.orbit;
  return 1;
}

void oracle_triplets(void* h, long* rows, long* cols, double* re, double* im) {
  auto* m = static_cast<Model*>(h);
  for (size_t i = 0; i < m->entries.size(); ++i) {
    rows[i] = m->entries[i].row;
    cols[i] = m->entries[i].col;
    re[i] = m->entries[i].value.real();
    im[i] = m->entries[i].value.imag();
  }
}

// complex interleaved vectors
void oracle_apply(void* h, const double* x, double* y) {
  auto* m = static_cast<Model*>(h);
  apply(*m, reinterpret_cast<const cplx*>(x), reinterpret_cast<cplx*>(y));
}
void oracle_apply_real(void* h, const double* x, double* y) { apply_real(*static_cast<Model*>(h), x, y); }

void oracle_deterministic_start(long n, unsigned long long seed, double* out) {
  auto v = deterministic_start(n, seed);
  std::copy(v.begin(), v.end(), out);
}

// eigsh: values/residuals sized k; vectors (k*n complex interleaved) optional;
// info[0]=iterations info[1]=reorth passes info[2]=second passes
int oracle_eigsh(void* h, int k, int target, long krylov, double tol, long max_iter, unsigned long long seed,
                 const double* initial, double* values, double* residuals, double* vectors, long* info) {
  auto* m = static_cast<Model*>(h);
  try {
    EigshOptions o{k, target, krylov, tol, max_iter, seed};
    std::vector<cplx> init;
    if (initial) init.assign(reinterpret_cast<const cplx*>(initial), reinterpret_cast<const cplx*>(initial) + m->dim());
    Eigensystem es = eigsh(*m, o, initial ? &init : nullptr);
    for (int i = 0; i < k; ++i) {
      values[i] = es.values[i];
      residuals[i] = es.residuals[i];
      if (vectors) std::copy(es.vectors[i].begin(), es.vectors[i].end(), reinterpret_cast<cplx*>(vectors) + size_t(i) * m->dim());
    }
    info[0] = es.iterations;
    info[1] = es.passes;
    info[2] = es.second;
    return 0;
  } catch (NonConvergence& nc) {
    std::ostringstream os;
    os << "solver did not converge after " << nc.iterations << " iterations; residual=" << nc.residual;
    g_err = os.str();
    return -2;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// eigsh from an initial subspace of nsub complex interleaved vectors
// (eigsh_with_initial_subspace, mode 0) or a workspace holding them
// (eigsh_with_workspace, mode 1); outputs as oracle_eigsh
int oracle_eigsh_subspace(void* h, int k, int target, long krylov, double tol, long max_iter, unsigned long long seed,
                          int nsub, const double* subspace, int mode, double* values, double* residuals,
                          double* vectors, long* info) {
  auto* m = static_cast<Model*>(h);
  try {
    EigshOptions o{k, target, krylov, tol, max_iter, seed};
    std::vector<std::vector<cplx>> sub;
    for (int j = 0; j < nsub; ++j) {
      const cplx* v = reinterpret_cast<const cplx*>(subspace) + size_t(j) * m->dim();
      sub.emplace_back(v, v + m->dim());
    }
    Eigensystem es = mode == 1 ? eigsh_workspace(*m, o, sub) : eigsh_subspace(*m, o, sub);
    for (int i = 0; i < k; ++i) {
      values[i] = es.values[i];
      residuals[i] = es.residuals[i];
      if (vectors) std::copy(es.vectors[i].begin(), es.vectors[i].end(), reinterpret_cast<cplx*>(vectors) + size_t(i) * m->dim());
    }
    info[0] = es.iterations;
    info[1] = es.passes;
    info[2] = es.second;
    return 0;
  } catch (NonConvergence& nc) {
    std::ostringstream os;
    os << "solver did not converge after " << nc.iterations << " iterations; residual=" << nc.residual;
    g_err = os.str();
    return -2;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// balanced_subspace_start of nsub complex interleaved vectors of length n
int oracle_balanced_start(int nsub, long n, const double* subspace, double* out) {
  try {
    std::vector<std::vector<cplx>> sub;
    for (int j = 0; j < nsub; ++j) {
      const cplx* v = reinterpret_cast<const cplx*>(subspace) + size_t(j) * n;
      sub.emplace_back(v, v + n);
    }
    auto r = balanced_subspace_start(sub);
    std::copy(r.begin(), r.end(), reinterpret_cast<cplx*>(out));
    return 0;
  } catch (NonConvergence&) {
    g_err = "initial subspace combines to zero";
    return -2;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// evolve one column; out is n_times*n complex interleaved; stats[5]
int oracle_evolve(void* h, const double* psi, const double* times, int n_times, long krylov, double tol,
                  long max_substeps, double* out, double* stats) {
  auto* m = static_cast<Model*>(h);
  try {
    std::vector<cplx> init(reinterpret_cast<const cplx*>(psi), reinterpret_cast<const cplx*>(psi) + m->dim());
    std::vector<double> ts(times, times + n_times);
    EvolveStats st;
    auto states = evolve(*m, init, ts, krylov, tol, max_substeps, st);
    for (int t = 0; t < n_times; ++t)
      std::copy(states[t].begin(), states[t].end(), reinterpret_cast<cplx*>(out) + size_t(t) * m->dim());
    if (stats) {
      stats[0] = double(st.projections);
      stats[1] = double(st.matvecs);
      stats[2] = double(st.accepted);
      stats[3] = double(st.rejected);
      stats[4] = st.max_error;
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Taylor plan: info = {degree, scaling}, dinfo = {shift_re, shift_im, one_norm};
// applies to n_vec columns (complex interleaved, column-major) in place.
int oracle_expm(void* h, double coeff_re, double coeff_im, int max_degree, double tol, long max_substeps,
                int n_vec, double* vecs, long* info, double* dinfo) {
  auto* m = static_cast<Model*>(h);
  try {
    Plan p = make_plan(*m, cplx(coeff_re, coeff_im), max_degree, tol, max_substeps);
    info[0] = p.degree;
    info[1] = p.scaling;
    dinfo[0] = p.shift.real();
    dinfo[1] = p.shift.imag();
    dinfo[2] = p.one_norm;
    for (int c = 0; c < n_vec; ++c) {
      cplx* col = reinterpret_cast<cplx*>(vecs) + size_t(c) * m->dim();
      std::vector<cplx> v(col, col + m->dim());
      auto r = plan_apply(*m, p, v);
      std::copy(r.begin(), r.end(), col);
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// dense Hermitian eigenvalues (row-major complex interleaved), for pinning
void oracle_hermitian_eigvals(int n, const double* a, double* vals) {
  std::vector<cplx> m(reinterpret_cast<const cplx*>(a), reinterpret_cast<const cplx*>(a) + size_t(n) * n);
  std::vector<double> v;
  std::vector<cplx> vec;
  hermitian_eig(n, m, v, vec);
  std::copy(v.begin(), v.end(), vals);
}

}  // extern "C"
