// Host-side compilation of local operator products into flip groups, plus the
// assembly-time checks of OperatorBuilder::build (QMBED
// src/operator/mod.rs:1941-2216).
//
// For spin one-half every local product maps a basis state to at most one
// state, and the set of flipped bits does not depend on the state: '+', '-',
// 'x', 'y' flip their site, 'z' and 'I' do not. Couplings that flip the same
// bits are merged into one FlipGroup whose amplitude table, indexed by the
// bits of the source state at the group's sites, holds c·a — the column
// entries of the reference's column-streaming assembly. For a Hermitian
// sector matrix the kernels read row r as the conjugate of column r
// (gather form).
#include <algorithm>
#include <cmath>
#include <map>
#include <sstream>
#include <tuple>

#include "internal.hpp"

namespace qmbed {

static const double F64_EPS = 2.220446049250313e-16;

static std::string strip_bars(const std::string& s) {
  std::string o;
  for (char c : s)
    if (c != '|') o += c;
  return o;
}

void validate_terms(const std::vector<OperatorSpec>& terms, int /*sites*/) {
  for (const auto& t : terms) {
    std::string sym = strip_bars(t.symbols);
    for (const auto& c : t.couplings) {
      if (c.sites.size() != sym.size()) {
        std::ostringstream os;
        os << "operator \"" << t.symbols << "\" has arity " << sym.size() << ", but a coupling has "
           << c.sites.size() << " sites";
        throw invalid_coupling(os.str());
      }
      if (!std::isfinite(c.coefficient.real()) || !std::isfinite(c.coefficient.imag()))
        throw invalid_coupling("coupling coefficients must be finite");
    }
  }
}

// Per-digit factors of SpinBasisBuilder::build for spin one-half
// (QMBED src/basis/mod.rs:1348-1376).
struct SpinFactors {
  double z[2], raise[2], lower[2], xy;
  explicit SpinFactors(Normalization n) {
    double zs = (n == Normalization::AngularMomentum) ? 1.0 : 2.0;
    z[0] = -0.5 * zs;
    z[1] = 0.5 * zs;
    double ladder = (n == Normalization::Pauli) ? 2.0 : 1.0;
    raise[0] = ladder;
    raise[1] = 0.0;
    lower[0] = 0.0;
    lower[1] = ladder;
    xy = (n == Normalization::PauliCartesian) ? 1.0 : 0.5;
  }
};

// Apply symbols right to left on a small packed state (bit i <-> local slot i),
// QMBED src/basis/mod.rs:1122-1232. Returns false when the action vanishes.
static bool local_apply(const SpinFactors& f, const std::string& sym, const std::vector<int>& slot, uint32_t in,
                        uint32_t& out, cplx& amp) {
  amp = cplx(1.0, 0.0);
  uint32_t s = in;
  for (int pos = int(sym.size()) - 1; pos >= 0; --pos) {
    int b = slot[pos];
    int digit = (s >> b) & 1;
    switch (sym[pos]) {
      case 'I': break;
      case 'z': amp *= f.z[digit]; break;
      case '+':
        if (f.raise[digit] == 0.0) return false;
        s ^= 1u << b;
        if (f.raise[digit] != 1.0) amp *= f.raise[digit];
        break;
      case '-':
        if (f.lower[digit] == 0.0) return false;
        s ^= 1u << b;
        if (f.lower[digit] != 1.0) amp *= f.lower[digit];
        break;
      case 'x':
      case 'y': {
        bool y = sym[pos] == 'y';
        if (f.raise[digit] != 0.0) {
          amp *= (y ? cplx(0, -f.xy) : cplx(f.xy, 0)) * f.raise[digit];
        } else {
          amp *= (y ? cplx(0, f.xy) : cplx(f.xy, 0)) * f.lower[digit];
        }
        s ^= 1u << b;
        break;
      }
      default: throw invalid_operator(std::string(1, sym[pos]));
    }
  }
  out = s;
  return true;
}

// operator_sum_preserves_particle_sector with the spin basis' sector test
// (QMBED src/operator/mod.rs:1963-2033, src/basis/mod.rs:294-371).
static bool preserves_particle_sector(const SpinBasisSpec& spec, const std::vector<OperatorSpec>& terms) {
  using Key = std::tuple<int, std::string, std::vector<int>>;
  std::map<Key, cplx> comps;
  for (const auto& t : terms) {
    for (const auto& c : t.couplings) {
      std::vector<std::tuple<int, std::string, cplx>> parts{{0, std::string(), cplx(1, 0)}};
      for (char s : t.symbols) {
        if (s == 'x' || s == 'y') {
          std::vector<std::tuple<int, std::string, cplx>> next;
          for (auto& [d, p, co] : parts) {
            next.emplace_back(d + 1, p + '+', co * (s == 'x' ? cplx(1, 0) : cplx(0, -1)));
            next.emplace_back(d + 1, p + '-', co * (s == 'x' ? cplx(1, 0) : cplx(0, 1)));
          }
          parts.swap(next);
        } else if (s == 'I' || s == 'n' || s == 'z' || s == '+' || s == '-' || s == '|') {
          for (auto& pt : parts) std::get<1>(pt) += s;
        } else {
          throw invalid_operator(t.symbols);
        }
      }
      for (auto& [d, p, co] : parts) comps[Key{d, p, c.sites}] += c.coefficient * co;
    }
  }
  bool fixed = spec.up.has_value() && spec.up_sectors.empty();
  if (!fixed && spec.up_sectors.empty()) return true;
  const int maximum = spec.sites * spec.spin_twice;
  auto contains = [&](int sector) {
    if (!spec.up_sectors.empty())
      return std::binary_search(spec.up_sectors.begin(), spec.up_sectors.end(), sector);
    return sector == *spec.up;
  };
  for (const auto& kv : comps) {
    if (std::abs(kv.second) <= F64_EPS) continue;
    int change = 0;
    for (char s : std::get<1>(kv.first)) change += (s == '+') - (s == '-');
    auto ok = [&](int source) {
      int target = source + change;
      return target < 0 || target > maximum || contains(target);
    };
    if (!spec.up_sectors.empty()) {
      for (int s : spec.up_sectors)
        if (!ok(s)) return false;
    } else if (!ok(*spec.up)) {
      return false;
    }
  }
  return true;
}

// Group the couplings by flip mask; each table holds c·a(pattern) for the
// source state's bits, i.e. the column entries the reference's
// column-streaming assembly generates.
static void build_groups(const std::vector<OperatorSpec>& terms, const std::vector<int>& part_of, int nparts,
                         const SpinFactors& f, std::vector<FlipGroup>& diag, std::vector<FlipGroup>& off);

static void check_sites_and_symbols(const SpinBasisSpec& spec, const std::vector<OperatorSpec>& terms) {
  for (const auto& t : terms) {
    std::string sym = strip_bars(t.symbols);
    for (const auto& c : t.couplings) {
      for (int pos = int(sym.size()) - 1; pos >= 0; --pos) {
        int site = c.sites[pos];
        if (site < 0 || site >= spec.sites) throw invalid_site(site, spec.sites);
        char ch = sym[pos];
        if (ch != 'I' && ch != 'z' && ch != '+' && ch != '-' && ch != 'x' && ch != 'y')
          throw invalid_operator(std::string(1, ch));
      }
    }
  }
}

static bool groups_real(const std::vector<FlipGroup>& diag, const std::vector<FlipGroup>& off) {
  for (auto* v : {&diag, &off})
    for (auto& g : *v) {
      for (auto& a : g.amp)
        if (a.imag() != 0.0) return false;
      for (auto& t : g.parts)
        for (auto& a : t)
          if (a.imag() != 0.0) return false;
    }
  return true;
}

void compile_family_groups(const SpinBasisSpec& spec, const std::vector<std::vector<OperatorSpec>>& parts,
                           const AssemblyChecks& checks, bool nonempty, std::vector<FlipGroup>& diag,
                           std::vector<FlipGroup>& off, bool& real) {
  std::vector<OperatorSpec> all;
  std::vector<int> part_of;
  for (size_t p = 0; p < parts.size(); ++p)
    for (const auto& t : parts[p]) {
      all.push_back(t);
      part_of.push_back(int(p));
    }
  validate_terms(all, spec.sites);
  // the reference checks particle conservation on the static operator and on
  // every component separately (each is built with the model's checks)
  if (checks.particle_conservation)
    for (const auto& terms : parts)
      if (!preserves_particle_sector(spec, terms))
        throw invalid_sector("operator sum does not preserve the selected particle sector");
  SpinFactors f(spec.normalization);
  if (nonempty) check_sites_and_symbols(spec, all);
  build_groups(all, part_of, int(parts.size()), f, diag, off);
  real = groups_real(diag, off);
}

void compile_groups(const SpinBasisSpec& spec, const std::vector<OperatorSpec>& terms, const AssemblyChecks& checks,
                    bool nonempty, std::vector<FlipGroup>& diag, std::vector<FlipGroup>& off, bool& real) {
  validate_terms(terms, spec.sites);
  if (checks.particle_conservation && !preserves_particle_sector(spec, terms))
    throw invalid_sector("operator sum does not preserve the selected particle sector");
  SpinFactors f(spec.normalization);
  // symbol alphabet and sites are validated where the reference's local
  // action would meet them (only when the sector is non-empty)
  if (nonempty) check_sites_and_symbols(spec, terms);
  build_groups(terms, std::vector<int>(terms.size(), 0), 0, f, diag, off);
  real = groups_real(diag, off);
}

// nparts > 0 (families): couplings of term i belong to part part_of[i], and
// every group also keeps one table per part.
static void build_groups(const std::vector<OperatorSpec>& terms, const std::vector<int>& part_of, int nparts,
                         const SpinFactors& f, std::vector<FlipGroup>& diag, std::vector<FlipGroup>& off) {
  diag.clear();
  off.clear();
  std::map<std::vector<int>, size_t> diag_index;  // site set -> group
  for (size_t ti = 0; ti < terms.size(); ++ti) {
    const auto& t = terms[ti];
    const int part = part_of[ti];
    std::string sym = strip_bars(t.symbols);
    const std::string& adj = sym;
    for (const auto& c : t.couplings) {
      if (c.coefficient == cplx(0, 0)) continue;
      const std::vector<int>& sites = c.sites;
      std::vector<int> uniq = sites;
      std::sort(uniq.begin(), uniq.end());
      uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
      if (uniq.size() > 5) throw unsupported_backend("a local product may act on at most 5 distinct sites");
      std::vector<int> slot(sites.size());
      for (size_t i = 0; i < sites.size(); ++i)
        slot[i] = int(std::lower_bound(uniq.begin(), uniq.end(), sites[i]) - uniq.begin());
      // flip mask (pattern independent for spin one-half)
      uint32_t flip_local = 0;
      for (size_t i = 0; i < adj.size(); ++i)
        if (adj[i] == '+' || adj[i] == '-' || adj[i] == 'x' || adj[i] == 'y') flip_local ^= 1u << slot[i];
      uint64_t flip = 0;
      for (size_t u = 0; u < uniq.size(); ++u)
        if ((flip_local >> u) & 1) flip |= uint64_t(1) << uniq[u];
      // amplitudes per local pattern
      const int np = 1 << uniq.size();
      std::vector<cplx> local_amp(np, cplx(0, 0));
      for (int pat = 0; pat < np; ++pat) {
        uint32_t out;
        cplx a;
        if (local_apply(f, adj, slot, uint32_t(pat), out, a)) {
          local_amp[pat] = c.coefficient * a;
        }
      }
      // merge into a group
      FlipGroup* g = nullptr;
      if (flip == 0) {
        auto it = diag_index.find(uniq);
        if (it != diag_index.end()) g = &diag[it->second];
        else {
          diag.emplace_back();
          g = &diag.back();
          g->flip = 0;
          g->nsites = int(uniq.size());
          for (size_t u = 0; u < uniq.size(); ++u) g->sites[u] = uniq[u];
          g->amp.assign(size_t(1) << uniq.size(), cplx(0, 0));
          g->parts.assign(size_t(nparts), g->amp);
          diag_index[uniq] = diag.size() - 1;
        }
      } else {
        for (auto& cand : off) {
          if (cand.flip != flip) continue;
          std::vector<int> un(cand.sites, cand.sites + cand.nsites);
          un.insert(un.end(), uniq.begin(), uniq.end());
          std::sort(un.begin(), un.end());
          un.erase(std::unique(un.begin(), un.end()), un.end());
          if (un.size() > 5) continue;
          if (int(un.size()) != cand.nsites) {
            // widen the candidate's tables to the union of sites
            auto widen = [&](const std::vector<cplx>& narrow) {
              std::vector<cplx> wide(size_t(1) << un.size(), cplx(0, 0));
              for (size_t pat = 0; pat < wide.size(); ++pat) {
                int old = 0;
                for (int i = 0; i < cand.nsites; ++i) {
                  int pos = int(std::lower_bound(un.begin(), un.end(), cand.sites[i]) - un.begin());
                  old |= int((pat >> pos) & 1) << i;
                }
                wide[pat] = narrow[old];
              }
              return wide;
            };
            std::vector<cplx> wide = widen(cand.amp);
            for (auto& t : cand.parts) t = widen(t);
            cand.nsites = int(un.size());
            for (size_t u = 0; u < un.size(); ++u) cand.sites[u] = un[u];
            cand.amp.swap(wide);
          }
          g = &cand;
          break;
        }
        if (!g) {
          off.emplace_back();
          g = &off.back();
          g->flip = flip;
          g->nsites = int(uniq.size());
          for (size_t u = 0; u < uniq.size(); ++u) g->sites[u] = uniq[u];
          g->amp.assign(size_t(1) << uniq.size(), cplx(0, 0));
          g->parts.assign(size_t(nparts), g->amp);
        }
      }
      // accumulate this coupling's amplitudes into the group's table
      for (size_t pat = 0; pat < g->amp.size(); ++pat) {
        int lp = 0;
        for (size_t u = 0; u < uniq.size(); ++u) {
          int pos = 0;
          for (int i = 0; i < g->nsites; ++i)
            if (g->sites[i] == uniq[u]) pos = i;
          lp |= int((pat >> pos) & 1) << u;
        }
        g->amp[pat] += local_amp[lp];
        if (nparts > 0) g->parts[size_t(part)][pat] += local_amp[lp];
      }
    }
  }
  // drop groups whose amplitudes all vanish (in every part, for families)
  auto empty = [](const FlipGroup& g) {
    for (auto& a : g.amp)
      if (a != cplx(0, 0)) return false;
    for (auto& t : g.parts)
      for (auto& a : t)
        if (a != cplx(0, 0)) return false;
    return true;
  };
  diag.erase(std::remove_if(diag.begin(), diag.end(), empty), diag.end());
  off.erase(std::remove_if(off.begin(), off.end(), empty), off.end());
}

}  // namespace qmbed
