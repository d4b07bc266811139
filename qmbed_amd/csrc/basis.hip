// Spin-1/2 sector enumeration on gfx950.
//
// Restates SpinBasisBuilder::build + spin_symmetry_sector
// (QMBED src/basis/mod.rs:373-440, 715-848, 1312-1412) as a two-pass
// predicate + ordered stream compaction over the parent space:
//   parent states (Gosper/colex order for a fixed up count, plain integers
//   otherwise) -> keep s iff s is the minimum of its translation/reflection
//   orbit and the orbit survives the character projection -> ascending reps.
// The character sums of the reference, Σ_s χ(s) over an orbit of period d,
// vanish exactly unless k·d ≡ 0 (mod L) (and, with parity, unless
// p·χ(m) = 1 where P r = T^m r); the predicate evaluates these conditions in
// integers, so the surviving set equals the reference's (|c| > 1e-12 test)
// without floating-point phases. The orbit size is the number of distinct
// states carrying a nonzero coefficient: d, or 2d when the reflection maps the
// translation orbit to a different one.
//
// General symmetry sectors (GeneralBasis::from_orbit_seeds_with_ordering +
// SymmetryReducer, src/basis/mod.rs:3464-3541, 4163-4300, built by the C
// ABI's build_spin_basis, bindings/capi/src/lib.rs:2849-2920): the group G
// generated by the maps is closed on the host, every element being
// s -> P(s) XOR f (a site permutation and a digit flip mask), with its
// character χ(g) as an integer index modulo the lcm of the generator
// periods. The reference's breadth-first orbit trace accepts an orbit iff a
// phase function c(g s) = χ(g) c(s) exists on it; that holds for every orbit
// or for none when two words for one element carry different phases (then
// the sector is empty), and otherwise iff χ(h) = 1 for every h fixing the
// seed. Each parent state s is the seed of its orbit iff no element maps it
// to a smaller parent state; the representative is the orbit's minimum (or
// maximum under the reverse convention) and the orbit size |G| / |Stab(s)|.
// The representatives are then sorted (the parent's ascending rows).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "qmbed.hpp"

namespace qmbed {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 64;       // parents per thread

struct EnumParams {
  int L;
  uint64_t mask;
  int up;                // -1: enumerate 0..2^L-1
  uint64_t popmask;      // allowed popcounts when enumerating 0..2^L-1 (bit i: popcount i)
  int kmom;              // -1: no translation symmetry
  int parity;            // 0 none
  uint64_t n_parents;
  const uint64_t* binom; // binom[n*65+k], n, k <= 64
};

// States are 32-bit words for L <= 31 and 64-bit words (wide sectors) for
// 32 <= L <= 63; these helpers are written once for both widths.
template <typename ST>
__device__ __forceinline__ int popc_st(ST s) {
  if constexpr (sizeof(ST) == 4) return __popc(s);
  else return __popcll(s);
}

// s rotated left by j (0 <= j < L) inside L bits; a narrow state rotates
// inside its doubled 64-bit word y = s | s << L
template <typename ST>
__device__ __forceinline__ ST rot_st(ST s, int j, int L, uint64_t mask) {
  if constexpr (sizeof(ST) == 4) {
    const uint64_t y = uint64_t(s) | (uint64_t(s) << L);
    return uint32_t(y >> (L - j)) & uint32_t(mask);
  } else {
    return j ? (((s << j) | (s >> (L - j))) & mask) : s;
  }
}

template <typename ST>
__device__ __forceinline__ ST reflect_st(ST s, int L) {
  if constexpr (sizeof(ST) == 4) return __brev(s) >> (32 - L);
  else return __brevll(s) >> (64 - L);
}

template <typename ST>
__device__ ST unrank_colex(uint64_t rank, int L, int up, const uint64_t* binom) {
  ST s = 0;
  int hi = L - 1;
  for (int k = up; k >= 1; --k) {
    int lo = k - 1;
    int h = hi;
    while (lo < h) {
      int mid = (lo + h + 1) >> 1;
      if (binom[mid * 65 + k] <= rank) lo = mid;
      else h = mid - 1;
    }
    s |= ST(1) << lo;
    rank -= binom[lo * 65 + k];
    hi = lo - 1;
  }
  return s;
}

template <typename ST>
__device__ __forceinline__ ST gosper_next(ST s) {
  ST low = s & (ST(0) - s);
  ST ripple = s + low;
  if constexpr (sizeof(ST) == 4) return (((ripple ^ s) >> 2) / low) | ripple;
  else return (((ripple ^ s) >> 2) >> __ffsll(static_cast<long long>(low)) - 1) | ripple;  // low is a power of two
}

// Returns the orbit size (>0) if s is a kept representative, 0 otherwise.
template <typename ST>
__device__ int rep_orbit(ST s, const EnumParams& p) {
  const int L = p.L;
  if (p.kmom < 0 && p.parity == 0) return 1;
  if (p.kmom < 0) {  // reflection only: orbit {s, P s}
    ST ps = reflect_st(s, L);
    if (ps < s) return 0;
    if (ps == s) return p.parity == 1 ? 1 : 0;
    return 2;
  }
  int d = L;
  for (int j = 1; j < L; ++j) {
    ST r = rot_st(s, j, L, p.mask);
    if (r < s) return 0;
    if (r == s && d == L) d = j;
  }
  if ((p.kmom * d) % L != 0) return 0;
  if (p.parity == 0) return d;
  ST ps = reflect_st(s, L);
  int jm = -1;
  for (int j = 0; j < L; ++j) {
    ST r = rot_st(ps, j, L, p.mask);
    if (r < s) return 0;
    if (r == s && jm < 0) jm = j;
  }
  if (jm < 0) return 2 * d;
  int m = (L - jm) % L;  // P s = T^m s
  int chi = (p.kmom == 0) ? 1 : ((m & 1) ? -1 : 1);
  return (p.parity * chi == 1) ? d : 0;
}

template <typename ST, bool WRITE>
__global__ void __launch_bounds__(kThreads) enumerate_kernel(EnumParams p, uint32_t* block_counts,
                                                             const uint64_t* block_offsets, ST* states,
                                                             uint8_t* orbits) {
  __shared__ uint32_t scan[kThreads];
  const uint64_t tid = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
  const uint64_t first = tid * kChunk;
  uint32_t count = 0;
  ST s = 0;
  int todo = 0;
  if (first < p.n_parents) {
    todo = int(min<uint64_t>(kChunk, p.n_parents - first));
    s = (p.up >= 0) ? unrank_colex<ST>(first, p.L, p.up, p.binom) : ST(first);
    ST t = s;
    for (int i = 0; i < todo; ++i) {
      bool ok = (p.up >= 0) || ((p.popmask >> popc_st(t)) & 1ull);
      if (ok && rep_orbit(t, p) > 0) ++count;
      if (i + 1 < todo) t = (p.up >= 0) ? gosper_next(t) : t + 1;
    }
  }
  scan[threadIdx.x] = count;
  __syncthreads();
  // inclusive Hillis-Steele scan over the block
  for (int off = 1; off < kThreads; off <<= 1) {
    uint32_t v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  if (!WRITE) {
    if (threadIdx.x == kThreads - 1) block_counts[blockIdx.x] = scan[kThreads - 1];
    return;
  }
  uint64_t out = block_offsets[blockIdx.x] + scan[threadIdx.x] - count;
  ST t = s;
  for (int i = 0; i < todo; ++i) {
    bool ok = (p.up >= 0) || ((p.popmask >> popc_st(t)) & 1ull);
    if (ok) {
      int o = rep_orbit(t, p);
      if (o > 0) {
        states[out] = t;
        orbits[out] = uint8_t(o);
        ++out;
      }
    }
    if (i + 1 < todo) t = (p.up >= 0) ? gosper_next(t) : t + 1;
  }
}

// g(s) for a general group element: the bit of site i moves to dest[i], then
// the flipped digits toggle
template <typename ST>
__device__ __forceinline__ ST group_apply(ST s, const uint8_t* __restrict__ dest, uint64_t flip, int L) {
  ST out = 0;
  for (int i = 0; i < L; ++i) out |= ((s >> i) & ST(1)) << dest[i];
  return out ^ ST(flip);
}

struct GroupParams {
  const uint8_t* dest;   // [n][64]
  const uint64_t* flip;  // [n]
  const uint8_t* phase;  // [n]: character index of χ(g⁻¹) (0 iff χ(g) = 1)
  // [n] when every element moves the sites as a rotation or a reflection of
  // the ring: k (the state rotated left by k) or (k | 0x40: its bit
  // reversal rotated left by k); null otherwise
  const uint8_t* rot;
  int n;
  int max_rep;
};

// s rotated left by k inside L bits (0 <= k < L)
template <typename ST>
__device__ __forceinline__ ST rotl_sites(ST s, int k, int L, uint64_t mask) {
  return k ? ST(((s << k) | (s >> (L - k))) & ST(mask)) : s;
}

// g(s) for an element of a rotation/reflection group: sr is s's bit reversal
// inside L bits
template <typename ST>
__device__ __forceinline__ ST group_apply_ring(ST s, ST sr, uint8_t code, uint64_t flip, int L, uint64_t mask) {
  return rotl_sites<ST>((code & 0x40) ? sr : s, code & 63, L, mask) ^ ST(flip);
}

template <typename ST>
__device__ __forceinline__ bool in_parent(ST t, const EnumParams& p) {
  return p.up >= 0 ? popc_st(t) == p.up : ((p.popmask >> popc_st(t)) & 1ull) != 0;
}

// s seeds a compatible orbit: returns its orbit size and representative
// (0 when s is not the seed or the orbit does not survive the projection)
template <typename ST>
__device__ int general_rep_orbit(ST s, const EnumParams& p, const GroupParams& gp, ST& rep) {
  rep = s;
  int stab = 0;
  bool ok = true;
  const ST sr = gp.rot ? reflect_st(s, p.L) : ST(0);
  for (int g = 0; g < gp.n; ++g) {
    const ST t = gp.rot ? group_apply_ring<ST>(s, sr, gp.rot[g], gp.flip[g], p.L, p.mask)
                        : group_apply(s, gp.dest + 64 * g, gp.flip[g], p.L);
    if (t == s) {
      ++stab;
      if (gp.phase[g] != 0) ok = false;
    } else if (t < s && in_parent(t, p)) {
      ok = false;  // a smaller parent state seeds this orbit
    }
    rep = gp.max_rep ? max(rep, t) : min(rep, t);
  }
  return ok ? gp.n / stab : 0;
}

template <typename ST, bool WRITE>
__global__ void __launch_bounds__(kThreads) general_enumerate_kernel(EnumParams p, GroupParams gp,
                                                                     uint32_t* block_counts,
                                                                     const uint64_t* block_offsets, ST* states,
                                                                     uint8_t* orbits) {
  __shared__ uint32_t scan[kThreads];
  const uint64_t tid = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
  const uint64_t first = tid * kChunk;
  uint32_t count = 0;
  ST s = 0;
  int todo = 0;
  if (first < p.n_parents) {
    todo = int(min<uint64_t>(kChunk, p.n_parents - first));
    s = (p.up >= 0) ? unrank_colex<ST>(first, p.L, p.up, p.binom) : ST(first);
    ST t = s, rep;
    for (int i = 0; i < todo; ++i) {
      if (in_parent(t, p) && general_rep_orbit(t, p, gp, rep) > 0) ++count;
      if (i + 1 < todo) t = (p.up >= 0) ? gosper_next(t) : t + 1;
    }
  }
  scan[threadIdx.x] = count;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    uint32_t v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  if (!WRITE) {
    if (threadIdx.x == kThreads - 1) block_counts[blockIdx.x] = scan[kThreads - 1];
    return;
  }
  uint64_t out = block_offsets[blockIdx.x] + scan[threadIdx.x] - count;
  ST t = s, rep;
  for (int i = 0; i < todo; ++i) {
    if (in_parent(t, p)) {
      const int o = general_rep_orbit(t, p, gp, rep);
      if (o > 0) {
        states[out] = rep;
        orbits[out] = uint8_t(o);
        ++out;
      }
    }
    if (i + 1 < todo) t = (p.up >= 0) ? gosper_next(t) : t + 1;
  }
}

// rows in descending order: out[i] = in[n - 1 - i]
template <typename T>
__global__ void reverse_kernel(const T* __restrict__ in, uint64_t n, T* __restrict__ out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[n - 1 - i];
}

// First index in [lo, hi) of the sorted states whose value is >= v (hi
// when none is).
__device__ uint32_t lower_bound_state(const uint32_t* __restrict__ states, uint32_t lo, uint32_t hi, uint64_t v) {
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (uint64_t(states[mid]) < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// State directory, pass 1: start[c] = first state of cell c (states whose
// value >> cb is c), for c = 0 .. cells (start[cells] = n).
__global__ void dir_start_kernel(const uint32_t* __restrict__ states, uint32_t n, int cb, uint32_t cells,
                                 uint32_t* __restrict__ start) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > cells) return;
  start[c] = lower_bound_state(states, 0, n, uint64_t(c) << cb);
}

// Pass 2: per cell the largest shift s <= cb at which no sub-bucket
// (states sharing value >> s) holds more than k states. The longest run of
// equal value >> s only shrinks as s falls, so s is found by bisection
// over passes through the cell's states; s = 0 always qualifies (the
// states are distinct).
__global__ void dir_shift_kernel(const uint32_t* __restrict__ states, const uint32_t* __restrict__ start, int cb,
                                 uint32_t cells, uint32_t k, uint32_t* __restrict__ shift) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const uint32_t lo = start[c], hi = start[c + 1];
  auto fits = [&](int s) {
    uint32_t run = 0, prev = 0;
    for (uint32_t i = lo; i < hi; ++i) {
      const uint32_t v = states[i] >> s;
      run = (i > lo && v == prev) ? run + 1 : 1;
      if (run > k) return false;
      prev = v;
    }
    return true;
  };
  int good = 0, bad = cb + 1;  // fits(good), and every s >= bad fails
  while (bad - good > 1) {
    const int mid = (good + bad) / 2;
    if (fits(mid)) good = mid;
    else bad = mid;
  }
  shift[c] = uint32_t(good);
}

// Pass 3: slot j (4 words) of the directory: the index of the first state
// of sub-bucket j and its states (at most 3, ascending, then all ones). The
// cell of j comes from a bisection over the cells' first slots.
__global__ void dir_fill_kernel(const uint32_t* __restrict__ states, const uint32_t* __restrict__ start,
                                const uint32_t* __restrict__ shift, const uint64_t* __restrict__ base, int cb,
                                uint32_t cells, uint64_t total, uint32_t* __restrict__ slots) {
  const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= total) return;
  uint32_t c = 0, hi = cells;  // largest c with base[c] <= j
  while (hi - c > 1) {
    const uint32_t mid = (c + hi) / 2;
    if (base[mid] <= j) c = mid;
    else hi = mid;
  }
  const uint64_t local = j - base[c];
  const uint64_t v = (uint64_t(c) << cb) + (local << shift[c]);
  const uint64_t vend = (uint64_t(c) << cb) + ((local + 1) << shift[c]);
  const uint32_t first = lower_bound_state(states, start[c], start[c + 1], v);
  const uint32_t last = lower_bound_state(states, first, start[c + 1], vend);
  uint32_t* out = slots + 4 * j;
  out[0] = first;
  for (uint32_t i = 0; i < 3; ++i) out[1 + i] = first + i < last ? states[first + i] : 0xFFFFFFFFu;
}

uint64_t binom_host(int n, int k) {
  if (k < 0 || k > n) return 0;
  unsigned __int128 r = 1;
  for (int i = 1; i <= k; ++i) r = r * unsigned(n - k + i) / unsigned(i);
  return uint64_t(r);
}

// The symmetry group of a general sector, closed on the host.
struct GroupElement {
  std::array<uint8_t, 64> dest{};
  uint64_t flip = 0;
  long chi = 0;  // χ(g) = exp(-2πi chi / M)
};

struct Group {
  std::vector<GroupElement> elems;
  long M = 1;
  bool consistent = true;  // one phase per element
};

constexpr int kMaxGroup = 255;  // orbit sizes are stored in 8 bits
constexpr long kMaxCharOrder = 1L << 20;
constexpr int kMaxChars = 64;   // character-table capacity of the H·v kernels

long gcd_l(long a, long b) {
  while (b) {
    const long t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// runtime_symmetry_sector + LatticeSymmetryMap::new for spin one-half
// (bindings/capi/src/lib.rs:3103-3127; src/basis/mod.rs:2978-3060,
// 3221-3268): map checks, generator periods, group closure
Group close_group(const std::vector<SymmetryMapSpec>& maps, int L) {
  struct Gen {
    std::array<uint8_t, 64> dest{};
    uint64_t flip = 0;  // destination-side flip mask
    long period = 1;
    long sector = 0;
  };
  std::vector<Gen> gens;
  for (const auto& m : maps) {
    if (long(m.destinations.size()) != long(L))
      throw invalid_options("symmetry map has " + std::to_string(m.destinations.size()) + " sites, expected " +
                            std::to_string(L));
    if (m.destinations.empty()) throw invalid_sector("a lattice symmetry requires sites and local states");
    std::vector<char> seen(size_t(L), 0);
    for (long d : m.destinations)
      if (d < 0 || d >= L || seen[size_t(d)]++)
        throw incompatible_symmetry("symmetry site destinations must be a bijection");
    std::vector<int> flipped(size_t(L), 0);
    if (!m.local_permutations.empty()) {
      if (long(m.local_permutations.size()) != long(L))
        throw incompatible_symmetry("every local-state map must be a permutation");
      for (int i = 0; i < L; ++i) {
        const auto& q = m.local_permutations[size_t(i)];
        const bool ident = q.size() == 2 && q[0] == 0 && q[1] == 1;
        const bool swap = q.size() == 2 && q[0] == 1 && q[1] == 0;
        if (!ident && !swap) throw incompatible_symmetry("every local-state map must be a permutation");
        flipped[size_t(i)] = swap ? 1 : 0;
      }
    }
    Gen g;
    for (int i = 0; i < L; ++i) {
      g.dest[size_t(i)] = uint8_t(m.destinations[size_t(i)]);
      if (flipped[size_t(i)]) g.flip |= uint64_t(1) << m.destinations[size_t(i)];
    }
    // combined_permutation_period: lcm of the cycles of (site, digit)
    std::vector<char> visited(size_t(2 * L), 0);
    long period = 1;
    for (int seed = 0; seed < 2 * L; ++seed) {
      if (visited[size_t(seed)]) continue;
      int cur = seed;
      long cycle = 0;
      while (!visited[size_t(cur)]) {
        visited[size_t(cur)] = 1;
        ++cycle;
        const int src = cur / 2, digit = cur % 2;
        cur = int(m.destinations[size_t(src)]) * 2 + (digit ^ flipped[size_t(src)]);
      }
      period = period / gcd_l(period, cycle) * cycle;
      if (period > kMaxCharOrder)
        throw unsupported_backend("symmetry map periods beyond 2^20 are not built on the MI355X path");
    }
    g.period = period;
    g.sector = ((long(m.sector) % period) + period) % period;
    gens.push_back(g);
  }
  Group G;
  for (auto& g : gens) {
    G.M = G.M / gcd_l(G.M, g.period) * g.period;
    if (G.M > kMaxCharOrder)
      throw unsupported_backend("symmetry character orders beyond 2^20 are not built on the MI355X path");
  }
  // breadth-first closure from the identity: h = gen ∘ e, i.e.
  // h(s) = P_gen(P_e(s) ^ f_e) ^ f_gen
  std::map<std::pair<std::array<uint8_t, 64>, uint64_t>, size_t> index;
  GroupElement id;
  for (int i = 0; i < L; ++i) id.dest[size_t(i)] = uint8_t(i);
  G.elems.push_back(id);
  index[{id.dest, id.flip}] = 0;
  for (size_t head = 0; head < G.elems.size(); ++head) {
    for (const auto& g : gens) {
      const GroupElement e = G.elems[head];
      GroupElement h;
      for (int i = 0; i < L; ++i) h.dest[size_t(i)] = g.dest[e.dest[size_t(i)]];
      uint64_t pf = 0;
      for (int i = 0; i < L; ++i)
        if ((e.flip >> i) & 1u) pf |= uint64_t(1) << g.dest[size_t(i)];
      h.flip = pf ^ g.flip;
      h.chi = (e.chi + g.sector * (G.M / g.period)) % G.M;
      auto it = index.find({h.dest, h.flip});
      if (it != index.end()) {
        if (G.elems[it->second].chi != h.chi) G.consistent = false;
        continue;
      }
      if (G.elems.size() >= size_t(kMaxGroup))
        throw unsupported_backend("general symmetry groups of more than 255 elements are not built on the MI355X path");
      index[{h.dest, h.flip}] = G.elems.size();
      G.elems.push_back(h);
    }
  }
  return G;
}

// Enumerate the sector's representatives into `states` (ascending, padded
// with all-ones words to whole 64-byte lines) and orbit sizes, and their
// reversal when rows descend; returns the dimension.
template <typename ST>
uint64_t enumerate_sector(EnumParams p, const GroupParams& gp, bool general, bool reversed, DeviceArray<ST>& states,
                          DeviceArray<ST>& rstates, DeviceArray<uint8_t>& orbits, DeviceArray<uint8_t>& rorbits,
                          hipStream_t stream) {
  const uint64_t threads = (p.n_parents + kChunk - 1) / kChunk;
  const uint64_t blocks = std::max<uint64_t>(1, (threads + kThreads - 1) / kThreads);
  if (blocks >= (uint64_t(1) << 31)) throw unsupported_backend("the parent sector is too large to enumerate");
  DeviceArray<uint32_t> d_counts(blocks);
  DeviceArray<uint64_t> d_offsets(blocks);
  if (general)
    hipLaunchKernelGGL((general_enumerate_kernel<ST, false>), dim3(unsigned(blocks)), dim3(kThreads), 0, stream, p,
                       gp, d_counts.data(), nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((enumerate_kernel<ST, false>), dim3(unsigned(blocks)), dim3(kThreads), 0, stream, p,
                       d_counts.data(), nullptr, nullptr, nullptr);
  QMBED_HIP(hipGetLastError());
  std::vector<uint32_t> counts(blocks);
  QMBED_HIP(hipMemcpyAsync(counts.data(), d_counts.data(), blocks * 4, hipMemcpyDeviceToHost, stream));
  QMBED_HIP(hipStreamSynchronize(stream));
  std::vector<uint64_t> offsets(blocks);
  uint64_t total = 0;
  for (uint64_t i = 0; i < blocks; ++i) {
    offsets[i] = total;
    total += counts[i];
  }
  const uint64_t line = 64 / sizeof(ST);
  const uint64_t padded = std::max<uint64_t>(line, (total + line - 1) / line * line);
  states.resize(padded);
  QMBED_HIP(hipMemsetAsync(states.data(), 0xFF, padded * sizeof(ST), stream));
  orbits.resize(std::max<uint64_t>(total, 1));
  QMBED_HIP(hipMemcpyAsync(d_offsets.data(), offsets.data(), blocks * 8, hipMemcpyHostToDevice, stream));
  if (general) {
    hipLaunchKernelGGL((general_enumerate_kernel<ST, true>), dim3(unsigned(blocks)), dim3(kThreads), 0, stream, p, gp,
                       nullptr, d_offsets.data(), states.data(), orbits.data());
    QMBED_HIP(hipGetLastError());
    // the representatives in ascending order (seeds are ascending, their
    // orbit maxima, or minima outside the parent, need not be)
    if (total > 1) {
      if (total >= (uint64_t(1) << 31)) throw unsupported_backend("general symmetry sector too large to sort");
      DeviceArray<ST> keys(total);
      DeviceArray<uint8_t> vals(total);
      QMBED_HIP(hipMemcpyAsync(keys.data(), states.data(), total * sizeof(ST), hipMemcpyDeviceToDevice, stream));
      QMBED_HIP(hipMemcpyAsync(vals.data(), orbits.data(), total, hipMemcpyDeviceToDevice, stream));
      size_t tmp_bytes = 0;
      QMBED_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys.data(), states.data(), vals.data(),
                                                   orbits.data(), int(total), 0, p.L, stream));
      DeviceArray<uint8_t> tmp(std::max<size_t>(tmp_bytes, 1));
      QMBED_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.data(), tmp_bytes, keys.data(), states.data(), vals.data(),
                                                   orbits.data(), int(total), 0, p.L, stream));
      QMBED_HIP(hipStreamSynchronize(stream));
    }
  } else {
    hipLaunchKernelGGL((enumerate_kernel<ST, true>), dim3(unsigned(blocks)), dim3(kThreads), 0, stream, p, nullptr,
                       d_offsets.data(), states.data(), orbits.data());
    QMBED_HIP(hipGetLastError());
  }
  if (reversed && total > 0) {
    rstates.resize(padded);
    QMBED_HIP(hipMemsetAsync(rstates.data(), 0xFF, padded * sizeof(ST), stream));
    rorbits.resize(total);
    const unsigned rb = unsigned((total + 255) / 256);
    hipLaunchKernelGGL(reverse_kernel<ST>, dim3(rb), dim3(256), 0, stream, states.data(), total, rstates.data());
    hipLaunchKernelGGL(reverse_kernel<uint8_t>, dim3(rb), dim3(256), 0, stream, orbits.data(), total, rorbits.data());
    QMBED_HIP(hipGetLastError());
  }
  return total;
}

}  // namespace

std::shared_ptr<SpinBasis1D> SpinBasis1D::build(const SpinBasisSpec& spec_in, hipStream_t stream) {
  SpinBasisSpec spec = spec_in;
  if (spec.spin_twice == 0) throw invalid_sector("spin_twice must be positive");
  if (spec.normalization != Normalization::AngularMomentum && spec.spin_twice != 1)
    throw invalid_options("the Pauli convention is defined only for spin one-half");
  if (!spec.symmetries.empty() && (spec.momentum || spec.parity))
    throw invalid_options("built-in and general spin symmetries cannot be mixed");
  if (spec.spin_twice != 1)
    throw unsupported_backend("the MI355X spin engine implements spin one-half sectors");
  const int L = spec.sites;
  if (L > 63) throw unsupported_backend("the MI355X spin engine packs states in 64 bits (at most 63 sites)");
  const bool wide = L > 31;  // 64-bit states
  if (spec.up && *spec.up > L) throw invalid_sector("particle count exceeds site count");
  for (int u : spec.up_sectors)
    if (u < 0 || u > L) throw invalid_sector("particle sector exceeds the spin capacity");
  std::sort(spec.up_sectors.begin(), spec.up_sectors.end());
  spec.up_sectors.erase(std::unique(spec.up_sectors.begin(), spec.up_sectors.end()), spec.up_sectors.end());
  if (spec.momentum || spec.parity) {
    if (L == 0) throw invalid_sector("symmetry sectors require at least one site");
  }
  if (spec.parity && *spec.parity != 1 && *spec.parity != -1)
    throw invalid_sector("parity must be either -1 or +1");
  int kmom = -1;
  if (spec.momentum) kmom = int(((long(*spec.momentum) % L) + L) % L);
  if (spec.parity && kmom > 0 && 2 * kmom != L)
    throw incompatible_symmetry("parity can share a one-dimensional sector with momentum only at k=0 or k=pi");
  if (L == 0) throw unsupported_backend("empty lattices are not supported by the MI355X spin engine");
  // the general sector's maps are checked and closed before any device work
  Group G;
  if (!spec.symmetries.empty()) G = close_group(spec.symmetries, L);

  auto b = std::shared_ptr<SpinBasis1D>(new SpinBasis1D());
  b->spec_ = spec;
  SectorGeometry& g = b->geo_;
  g.L = L;
  g.mask = (uint64_t(1) << L) - 1u;
  g.wide = wide;
  g.kmom = kmom;
  g.parity = spec.parity ? *spec.parity : 0;
  for (int d = 1; d < L; ++d)
    if (L % d == 0 && g.n_div < 16) g.divisors[g.n_div++] = d;

  EnumParams p{};
  p.L = L;
  p.mask = g.mask;
  p.kmom = kmom;
  p.parity = g.parity;
  if (spec.up && spec.up_sectors.empty()) {
    p.up = *spec.up;
    p.n_parents = binom_host(L, p.up);
  } else {
    if (L > 36) throw unsupported_backend("sectors without a fixed up count support at most 36 sites");
    p.up = -1;
    p.n_parents = uint64_t(1) << L;
    p.popmask = 0;
    if (spec.up_sectors.empty()) {
      p.popmask = ~0ull;
    } else {
      for (int u : spec.up_sectors) p.popmask |= 1ull << u;
    }
  }
  std::vector<uint64_t> binom(65 * 65);
  for (int n = 0; n <= 64; ++n)
    for (int k = 0; k <= 64; ++k) binom[n * 65 + k] = binom_host(n, k);
  DeviceArray<uint64_t> d_binom(binom.size());
  QMBED_HIP(hipMemcpyAsync(d_binom.data(), binom.data(), binom.size() * 8, hipMemcpyHostToDevice, stream));
  p.binom = d_binom.data();

  // general symmetry sector: the group, its character table and the
  // device copies of its elements
  GroupParams gp{};
  bool group_empty = false;
  if (!spec.symmetries.empty()) {
    group_empty = !G.consistent;
    const long M = G.M;
    // distinct values of χ(g⁻¹) = exp(-2πi (M - chi) / M), 1 first
    std::vector<long> vals;
    for (auto& e : G.elems) vals.push_back((M - e.chi) % M);
    std::sort(vals.begin(), vals.end());
    vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
    if (vals.size() > size_t(kMaxChars))
      throw unsupported_backend("general symmetry sectors with more than 64 character values are not built on the "
                                "MI355X path");
    const double tau = 6.283185307179586;
    b->chars_.clear();
    b->real_chars_ = true;
    for (long c : vals) {
      cplx v;
      if (c == 0) v = cplx(1, 0);
      else if (2 * c == M) v = cplx(-1, 0);
      else if (4 * c == M) v = cplx(0, -1);
      else if (4 * c == 3 * M) v = cplx(0, 1);
      else v = std::polar(1.0, -tau * double(c) / double(M));
      if (v.imag() != 0.0) b->real_chars_ = false;
      b->chars_.push_back(v);
    }
    const size_t n = G.elems.size();
    std::vector<uint8_t> dest(64 * n, 0), phase(n, 0);
    std::vector<uint64_t> flip(n, 0);
    for (size_t i = 0; i < n; ++i) {
      std::copy(G.elems[i].dest.begin(), G.elems[i].dest.end(), dest.begin() + long(64 * i));
      flip[i] = G.elems[i].flip;
      const long c = (M - G.elems[i].chi) % M;
      phase[i] = uint8_t(std::lower_bound(vals.begin(), vals.end(), c) - vals.begin());
    }
    // elements that move the sites as a rotation j -> j + k or a reflection
    // j -> k - j (mod L) of the ring (translations, parities and their
    // products, with any digit flips) apply as a rotation of the state or of
    // its bit reversal, a few instructions instead of one step per site
    std::vector<uint8_t> rot(n, 0);
    bool ring = true;
    for (size_t i = 0; i < n && ring; ++i) {
      const auto& d = G.elems[i].dest;
      const int k = int(d[0]);
      bool is_rot = true, is_ref = true;
      for (int j = 0; j < L; ++j) {
        if (int(d[size_t(j)]) != (j + k) % L) is_rot = false;
        if (int(d[size_t(j)]) != ((k - j) % L + L) % L) is_ref = false;
      }
      if (is_rot) rot[i] = uint8_t(k);
      else if (is_ref) rot[i] = uint8_t(((k + 1) % L) | 0x40);
      else ring = false;
    }
    if (ring) {
      b->grot_.resize(n);
      QMBED_HIP(hipMemcpyAsync(b->grot_.data(), rot.data(), n, hipMemcpyHostToDevice, stream));
    }
    b->gdest_.resize(dest.size());
    b->gflip_.resize(n);
    b->gphase_.resize(n);
    QMBED_HIP(hipMemcpyAsync(b->gdest_.data(), dest.data(), dest.size(), hipMemcpyHostToDevice, stream));
    QMBED_HIP(hipMemcpyAsync(b->gflip_.data(), flip.data(), n * 8, hipMemcpyHostToDevice, stream));
    QMBED_HIP(hipMemcpyAsync(b->gphase_.data(), phase.data(), n, hipMemcpyHostToDevice, stream));
    g.n_group = int(n);
    g.n_chars = int(vals.size());
    g.max_rep = spec.reverse;
    gp.dest = b->gdest_.data();
    gp.flip = b->gflip_.data();
    gp.phase = b->gphase_.data();
    gp.rot = b->group_rot();
    gp.n = int(n);
    gp.max_rep = spec.reverse ? 1 : 0;
  }
  g.reversed = spec.reverse;

  if (group_empty) p.n_parents = 0;  // every orbit trace conflicts: an empty sector
  uint64_t total = 0;
  if (wide) {
    total = enumerate_sector<uint64_t>(p, gp, g.n_group > 0, g.reversed, b->states64_, b->rstates64_, b->orbits_,
                                       b->rorbits_, stream);
  } else {
    total = enumerate_sector<uint32_t>(p, gp, g.n_group > 0, g.reversed, b->states_, b->rstates_, b->orbits_,
                                       b->rorbits_, stream);
  }
  b->n_ = total;
  if (wide) {
    // wide sectors are searched by a plain binary search (build-time
    // kernels only: their products run on the compiled stream)
    QMBED_HIP(hipStreamSynchronize(stream));
    return b;
  }
  build_directory(*b, stream);
  QMBED_HIP(hipStreamSynchronize(stream));
  return b;
}

// The state directory of a 32-bit sector (see SpinBasis1D::directory).
void SpinBasis1D::build_directory(SpinBasis1D& b, hipStream_t stream) {
  const uint32_t n = uint32_t(b.n_);
  const int L = b.geo_.L;
  const int bits = std::min(L, kDirCellBits);
  const int cb = L - bits;
  const uint32_t cells = 1u << bits;
  b.dir_shift_ = cb;
  DeviceArray<uint32_t> start(cells + 1), shift(cells);
  hipLaunchKernelGGL(dir_start_kernel, dim3((cells + 1 + 255) / 256), dim3(256), 0, stream, b.states_.data(), n, cb,
                     cells, start.data());
  QMBED_HIP(hipGetLastError());
  hipLaunchKernelGGL(dir_shift_kernel, dim3((cells + 63) / 64), dim3(64), 0, stream, b.states_.data(), start.data(),
                     cb, cells, uint32_t(kDirBucket), shift.data());
  QMBED_HIP(hipGetLastError());
  std::vector<uint32_t> sh(cells);
  QMBED_HIP(hipMemcpyAsync(sh.data(), shift.data(), size_t(cells) * 4, hipMemcpyDeviceToHost, stream));
  QMBED_HIP(hipStreamSynchronize(stream));
  std::vector<uint64_t> base(cells);
  std::vector<uint32_t> desc(2 * size_t(cells));
  uint64_t total = 0;
  for (uint32_t c = 0; c < cells; ++c) {
    base[c] = total;
    if (total >= (uint64_t(1) << 32)) throw unsupported_backend("sector too large for the state directory");
    desc[2 * size_t(c)] = uint32_t(total);
    desc[2 * size_t(c) + 1] = sh[c];
    total += uint64_t(1) << (cb - int(sh[c]));
  }
  if (total >= (uint64_t(1) << 32)) throw unsupported_backend("sector too large for the state directory");
  DeviceArray<uint64_t> dbase(cells);
  QMBED_HIP(hipMemcpyAsync(dbase.data(), base.data(), size_t(cells) * 8, hipMemcpyHostToDevice, stream));
  b.dir_cells_.resize(2 * size_t(cells));
  QMBED_HIP(hipMemcpyAsync(b.dir_cells_.data(), desc.data(), desc.size() * 4, hipMemcpyHostToDevice, stream));
  b.dir_.resize(4 * std::max<uint64_t>(total, 1));
  hipLaunchKernelGGL(dir_fill_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, stream, b.states_.data(),
                     start.data(), shift.data(), dbase.data(), cb, cells, total, b.dir_.data());
  QMBED_HIP(hipGetLastError());
  QMBED_HIP(hipStreamSynchronize(stream));
}

std::vector<uint32_t> SpinBasis1D::states_host() const {
  if (wide()) throw unsupported_backend("the sector's states are 64-bit words");
  std::vector<uint32_t> h(n_);
  if (n_) QMBED_HIP(hipMemcpy(h.data(), states_.data(), n_ * 4, hipMemcpyDeviceToHost));
  return h;
}

std::vector<uint64_t> SpinBasis1D::row_states_host64() const {
  std::vector<uint64_t> h(n_);
  if (wide()) {
    if (n_) QMBED_HIP(hipMemcpy(h.data(), states64_.data(), n_ * 8, hipMemcpyDeviceToHost));
  } else {
    const auto narrow = states_host();
    std::copy(narrow.begin(), narrow.end(), h.begin());
  }
  if (reversed()) std::reverse(h.begin(), h.end());
  return h;
}

std::vector<uint8_t> SpinBasis1D::orbits_host() const {
  std::vector<uint8_t> h(n_);
  if (n_) QMBED_HIP(hipMemcpy(h.data(), orbits_.data(), n_, hipMemcpyDeviceToHost));
  return h;
}

std::vector<uint32_t> SpinBasis1D::row_states_host() const {
  std::vector<uint32_t> h = states_host();
  if (reversed()) std::reverse(h.begin(), h.end());
  return h;
}

std::vector<uint8_t> SpinBasis1D::row_orbits_host() const {
  std::vector<uint8_t> h = orbits_host();
  if (reversed()) std::reverse(h.begin(), h.end());
  return h;
}

}  // namespace qmbed
