// Matrix-free, symmetry-reduced H·v for spin-1/2 chains on gfx950.
//
// One lane per output row (gather form, no atomics, deterministic order):
//   r   = representative of row `row`                  (coalesced u32 read)
//   y_r = Σ_diag-groups a(pattern(r)) · x_r
//       + Σ_{flip groups g active on r} a_g(pattern(r)) · sqrt(N_r/N_t) · χ(t) · x[idx(rep(t))]
// where t = r XOR flip_g. rep(t) and the group element mapping rep(t) to t are
// found by walking t's translation (and reflection) orbit in registers; the
// characters χ are tabulated in LDS; idx(rep) comes from the state
// directory: a cell table in LDS, one sub-bucket start (L2) and one window
// of the sorted representative table.
//
// This evaluates the same matrix as the reference's column-streaming
// assembly H[row(rep_t), col(r)] += c·a·sqrt(N_r/N_t)·conj(phase_t)
// (QMBED src/operator/mod.rs:2119-2205, src/basis/mod.rs:1561-1580), read
// row-wise: the row entries are the conjugated column entries of the adjoint
// product, which is what the flip-group tables hold.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <set>
#include <string>
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace qmbed {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxGroups = 256;  // a J1-J2 chain of 63 sites holds 252
constexpr int kMaxAmps = 1024;
constexpr int kMaxChars = 64;  // character-table entries staged in LDS

struct HvArgs {
  const uint32_t* states;   // ascending representatives (the search table)
  const uint32_t* rstates;  // representatives in row order
  const uint8_t* orbits;    // orbit sizes in row order
  int rev;                  // rows descending: row = n - 1 - ascending index
  // wide sectors (L >= 32): 64-bit representatives (ascending, and in row
  // order), the 64-bit lattice mask and per group [flip, sites (6 bits
  // each above a 3-bit count), nonzero-pattern mask, amplitude offset]
  const uint64_t* states64;
  const uint64_t* rstates64;
  uint64_t mask64;
  const uint64_t* meta64;
  uint32_t n;
  int L;
  uint32_t mask;
  int parity;
  int n_div;
  int divisors[16];
  // the state directory: per cell (first slot, shift) pairs (staged in LDS
  // by the walk-mode blocks), the 16-word slots of the sub-buckets, the cell
  // shift cb (cell = value >> cb) and the cell count
  const uint32_t* dcells;
  const uint32_t* dir;
  int dcb;
  int ncells;
  // lanes per row (power of two) and the mask of flip groups a lane owns
  int lpr;
  uint32_t lane_groups;
  // image-table mode: per parent-state rank, a packed (index, orbit code,
  // reflection sign, character index) word; 0xFFFFFFFF = not in the sector
  const uint32_t* img;
  int img_bits;          // bits of the representative index
  int up;                // fixed up count (colex ranks) or -1 (rank = state)
  const uint32_t* hop;   // per group: p | q << 8 | 1 << 16 for two-site hops
  const uint32_t* binom; // C(n, k) at [n * 32 + k], n, k < 32
  const double* code_rsq;  // 1/sqrt(orbit size) per orbit code (16)
  // groups (diag first, then off-diagonal)
  const uint32_t* meta;  // 4 words per group: flip, packed sites, nzmask, amp offset
  const double* amp;     // real: 1 double per entry; complex: 2 doubles per entry
  int n_diag, n_off, n_amp;
  int max_sites;         // the most sites of any group (at most 5)
  const double* chars;   // 2*nchar doubles (re, im): χ(a) = exp(-2πi k a / L), or the
                         // general sector's character values
  int nchar;
  // general symmetry group (ngroup elements; 0: built-in symmetries)
  const uint8_t* gdest;   // [ngroup][64] site destinations
  const uint64_t* gflip;  // destination-side flip masks
  const uint8_t* gphase;  // character index of χ(g⁻¹)
  // per element k or (k | 0x40) for a rotation/reflection group of the ring
  // (the state, or its bit reversal, rotated left by k), else null
  const uint8_t* grot;
  int ngroup;
  int maxrep;             // representatives are orbit maxima
  // batched columns
  int nvec;
  size_t ld;
  // optional device flag: a nonzero value makes the launch a no-op (a
  // speculative Lanczos step voided by a pending reorthogonalization)
  const int* gate;
  // compiled index stream (sliced ELL of packed target words), or null
  const uint32_t* stream;
  const uint32_t* slice_ptr;
  // compact stream words (compact != 0): idx | amplitude entry (relative to
  // off_base) << c_b | orbit code << (c_b + c_ob) | (sign < 0) << (c_b +
  // c_ob + c_cb) | χ-index << (c_b + c_ob + c_cb + c_sb); no group walk
  int compact;
  int c_b, c_ob, c_cb, c_sb, c_off_base;
  // split compact words (the fields do not fit 32 bits): the word holds
  // idx | entry << c_b, and one byte per word (same layout as the stream)
  // holds the fields above, orbit code | sign << c_cb | χ-index << (c_cb +
  // c_sb); null otherwise
  const uint8_t* aux;
  // per-row diagonal (re[n], then im[n]) of the current coefficients, or null
  const double* diag;
};

// `bits` wide field of e at `shift` (0 when bits == 0)
__device__ __forceinline__ uint32_t field(uint32_t e, int shift, int bits) {
  return bits ? (e >> shift) & ((1u << bits) - 1u) : 0u;
}

__device__ __forceinline__ uint32_t rotl_l(uint64_t y, int L, int j, uint32_t mask) {
  return uint32_t(y >> (L - j)) & mask;
}
__device__ __forceinline__ uint32_t reflect_l(uint32_t s, int L) { return __brev(s) >> (32 - L); }

// pattern_of with a uniform trip count (the operator's largest group): the
// loop is scalar-controlled, absent sites masked off at the end
__device__ __forceinline__ int pattern_upto(uint32_t r, uint32_t packed, int nmax) {
  int pat = 0;
  for (int i = 0; i < nmax; ++i) pat |= int(__builtin_amdgcn_ubfe(r, (packed >> (3 + 5 * i)) & 31u, 1)) << i;
  return pat & ((1 << (packed & 7u)) - 1);
}

__device__ __forceinline__ int pattern_of(uint32_t r, uint32_t packed) {
  const int ns = packed & 7;
  int pat = 0;
  for (int i = 0; i < ns; ++i) {
    int site = (packed >> (3 + 5 * i)) & 31;
    pat |= int((r >> site) & 1u) << i;
  }
  return pat;
}

// Index of each of up to B keys in the sorted state table (-1: absent, or
// in[b] false). The key's cell descriptor (LDS) names its sub-bucket
// j = first + (key mod 2^cb) >> shift, whose 16-byte slot holds the index
// of its first state and its states (at most 3, padded with all ones); the
// key is present exactly when one of them equals it. The slots of all B
// keys are loaded before any is used, unconditionally: every key below 2^L
// names a slot, and in[b] enters only the result.
template <int B>
__device__ __forceinline__ void dir_issue(const uint32_t (&key)[B], const uint32_t* cells, const HvArgs& a,
                                          uint4 (&w)[B]) {
  static_assert(SpinBasis1D::kDirBucket == 3, "a slot holds 3 states");
  const uint32_t lmask = (1u << a.dcb) - 1u;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const uint2 d = reinterpret_cast<const uint2*>(cells)[key[b] >> a.dcb];
    w[b] = __ldg(reinterpret_cast<const uint4*>(a.dir) + (d.x + ((key[b] & lmask) >> d.y)));
  }
}

template <int B>
__device__ __forceinline__ void dir_resolve(const uint32_t (&key)[B], const bool (&in)[B], const uint4 (&w)[B],
                                            int (&idx)[B]) {
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const uint32_t k = key[b];
    const int pos = w[b].y == k ? 0 : w[b].z == k ? 1 : w[b].w == k ? 2 : -1;
    idx[b] = in[b] && pos >= 0 ? int(w[b].x) + pos : -1;
  }
}

template <int B>
__device__ __forceinline__ void dir_find(const uint32_t (&key)[B], const bool (&in)[B], const uint32_t* cells,
                                         const HvArgs& a, int (&idx)[B]) {
  uint4 w[B];
  dir_issue<B>(key, cells, a, w);
  dir_resolve<B>(key, in, w, idx);
}

// dir_find for one key, with the cell table in global memory
__device__ __forceinline__ int dir_find1(uint32_t key, const HvArgs& a) {
  const uint32_t k[1] = {key};
  const bool in[1] = {true};
  int idx[1];
  dir_find<1>(k, in, a.dcells, a, idx);
  return idx[0];
}

// Colex (Gosper-order) rank of a fixed-weight state: Σ C(position, ordinal)
// over its set bits (QMBED src/basis/mod.rs:640-659).
__device__ __forceinline__ uint32_t rank_colex(uint32_t s, const uint32_t* binom) {
  uint32_t rank = 0;
  int ord = 1;
  while (s) {
    const int pos = __ffs(s) - 1;
    rank += binom[pos * 32 + ord];
    s &= s - 1;
    ++ord;
  }
  return rank;
}

// rank(s') - rank(s) when the set bit at p moves to the empty position q > p:
// bits strictly between p and q lose one ordinal, the moved bit gains them.
__device__ __forceinline__ int move_delta(uint32_t s, int p, int q, const uint32_t* binom) {
  const int a = __popc(s & ((1u << p) - 1u)) + 1;
  uint32_t mid = s & (((q >= 32) ? 0xFFFFFFFFu : ((1u << q) - 1u)) & ~((2u << p) - 1u));
  int m = __popc(mid);
  int d = int(binom[q * 32 + a + m]) - int(binom[p * 32 + a]);
  int k = a;
  while (mid) {
    const int b = __ffs(mid) - 1;
    d += int(binom[b * 32 + k]) - int(binom[b * 32 + k + 1]);
    mid &= mid - 1;
    ++k;
  }
  return d;
}

// Orbit walk of a parent state t, one rotation at a time (64-bit shift and
// mask per rotation); `walk` below computes the same and falls back to this
// form only for momentum phases when the rotation count does not fit below
// the state's bits (L >= 28).
// SYM: 1 translation, 2 reflection only, 3 translation + reflection.
// On return: rep, orbit size, and the phase as (ci: character index, sign)
// such that phase = sign * χ(ci).
template <int SYM, bool SHIFT>
__device__ __forceinline__ void walk_scan(uint32_t t, const HvArgs& a, uint32_t& rep, int& orbit, int& ci, int& sign) {
  const int L = a.L;
  sign = 1;
  ci = 0;
  if (SYM == 2) {
    uint32_t ps = reflect_l(t, L);
    if (ps < t) {
      rep = ps;
      orbit = 2;
      sign = a.parity;
    } else {
      rep = t;
      orbit = (ps == t) ? 1 : 2;
    }
    return;
  }
  const uint64_t y = uint64_t(t) | (uint64_t(t) << L);
  uint32_t mn = t;
  int jmin = 0;
  for (int j = 1; j < L; ++j) {
    uint32_t r = rotl_l(y, L, j, a.mask);
    if (SHIFT) {
      if (r < mn) {
        mn = r;
        jmin = j;
      }
    } else {
      mn = min(mn, r);
    }
  }
  int d = L;
#pragma unroll
  for (int i = 15; i >= 0; --i) {
    if (i < a.n_div) {
      const int dv = a.divisors[i];
      if (rotl_l(y, L, dv, a.mask) == t) d = dv;
    }
  }
  if (SYM == 1) {
    rep = mn;
    orbit = d;
    ci = (L - jmin) % L;
    return;
  }
  const uint32_t ps = reflect_l(t, L);
  const uint64_t yp = uint64_t(ps) | (uint64_t(ps) << L);
  uint32_t mnp = ps;
  int jminp = 0;
  for (int j = 1; j < L; ++j) {
    uint32_t r = rotl_l(yp, L, j, a.mask);
    if (SHIFT) {
      if (r < mnp) {
        mnp = r;
        jminp = j;
      }
    } else {
      mnp = min(mnp, r);
    }
  }
  if (mnp < mn) {
    rep = mnp;
    orbit = 2 * d;
    ci = (L - jminp) % L;
    sign = a.parity;
  } else {
    rep = mn;
    orbit = (mnp == mn) ? d : 2 * d;
    ci = (L - jmin) % L;
  }
}

// The orbit walk with one instruction per rotation: with the doubled state
// y = t·2^L + t left-aligned in 64 bits (hi:lo), v_alignbit(hi, lo, 32 − j)
// is a 32-bit word whose top L bits are rotl(t, j) and whose low 32 − L bits
// follow from the doubled string. The minimum over the words has rotl's
// minimum in its top bits whatever the low bits hold; with momentum phases
// the low bits carry j instead, so the minimum also names the first rotation
// that reaches it (the strict-less scan's choice). Same results as
// walk_scan for every state (checked exhaustively for L <= 20 and on
// samples and all periodic states up to L = 31).
template <int SYM, bool SHIFT>
__device__ __forceinline__ void walk(uint32_t t, const HvArgs& a, uint32_t& rep, int& orbit, int& ci, int& sign) {
  const int L = a.L;
  if (SYM == 2) {
    walk_scan<2, false>(t, a, rep, orbit, ci, sign);
    return;
  }
  const int lowb = 32 - L;  // L <= 31
  if (SHIFT && L > (1 << lowb)) {
    walk_scan<SYM, SHIFT>(t, a, rep, orbit, ci, sign);
    return;
  }
  sign = 1;
  const uint32_t low = (1u << lowb) - 1u;
  uint32_t hi, lo;
  auto min_word = [&](uint32_t s) {
    const uint64_t y = (uint64_t(s) | (uint64_t(s) << L)) << (64 - 2 * L);
    hi = uint32_t(y >> 32);
    lo = uint32_t(y);
    uint32_t mn = SHIFT ? (hi & ~low) : hi;
    auto word = [&](int j) {
      const uint32_t w = __builtin_amdgcn_alignbit(hi, lo, uint32_t(32 - j));
      return SHIFT ? (w & ~low) | uint32_t(j) : w;
    };
    // two rotations per step (one three-input minimum)
    int j = 1;
    for (; j + 1 < L; j += 2) mn = min(mn, min(word(j), word(j + 1)));
    if (j < L) mn = min(mn, word(j));
    return mn;
  };
  const uint32_t mn = min_word(t);
  // the orbit's period: the smallest divisor dv of L with rotl(t, dv) = t
  // (descending, the last match wins)
  int d = L;
#pragma unroll 1
  for (int i = a.n_div - 1; i >= 0; --i) {
    const int dv = a.divisors[i];
    if ((__builtin_amdgcn_alignbit(hi, lo, uint32_t(32 - dv)) >> lowb) == t) d = dv;
  }
  const uint32_t rep_o = mn >> lowb;
  const int j_o = SHIFT ? int(mn & low) : 0;
  if (SYM == 1) {
    rep = rep_o;
    orbit = d;
    ci = (L - j_o) % L;
    return;
  }
  const uint32_t mnp = min_word(reflect_l(t, L));
  const uint32_t rep_p = mnp >> lowb;
  const int j_p = SHIFT ? int(mnp & low) : 0;
  if (rep_p < rep_o) {
    rep = rep_p;
    orbit = 2 * d;
    ci = (L - j_p) % L;
    sign = a.parity;
  } else {
    rep = rep_o;
    orbit = (rep_p == rep_o) ? d : 2 * d;
    ci = (L - j_o) % L;
  }
}

// The orbit walks of B parent states at once, each as in `walk`, with the
// rotation and period loops shared by the B states so that the loop control
// (scalar instructions: L is a launch parameter) is paid once per step
// rather than once per state. Returns the representatives, orbit sizes and
// the phases as (ci, sign).
template <int SYM, bool SHIFT, int B>
__device__ __forceinline__ void walk_batch(const uint32_t (&t)[B], const HvArgs& a, uint32_t (&rep)[B], int (&orbit)[B],
                                           int (&ci)[B], int (&sign)[B]) {
  const int L = a.L;
  if (SYM == 2 || (SHIFT && L > (1 << (32 - L)))) {
#pragma unroll
    for (int b = 0; b < B; ++b) walk_scan<SYM, SHIFT>(t[b], a, rep[b], orbit[b], ci[b], sign[b]);
    return;
  }
  const int lowb = 32 - L;  // L <= 31
  const uint32_t low = (1u << lowb) - 1u;
  uint32_t hi[B], lo[B];
  auto min_words = [&](uint32_t (&mn)[B]) {
#pragma unroll
    for (int b = 0; b < B; ++b) mn[b] = SHIFT ? (hi[b] & ~low) : hi[b];
    auto word = [&](int b, int j) {
      const uint32_t w = __builtin_amdgcn_alignbit(hi[b], lo[b], uint32_t(32 - j));
      return SHIFT ? (w & ~low) | uint32_t(j) : w;
    };
    int j = 1;
    for (; j + 1 < L; j += 2) {
#pragma unroll
      for (int b = 0; b < B; ++b) mn[b] = min(mn[b], min(word(b, j), word(b, j + 1)));
    }
    if (j < L) {
#pragma unroll
      for (int b = 0; b < B; ++b) mn[b] = min(mn[b], word(b, j));
    }
  };
  auto doubled = [&](int b, uint32_t s) {
    const uint64_t y = (uint64_t(s) | (uint64_t(s) << L)) << (64 - 2 * L);
    hi[b] = uint32_t(y >> 32);
    lo[b] = uint32_t(y);
  };
  uint32_t mn[B];
#pragma unroll
  for (int b = 0; b < B; ++b) doubled(b, t[b]);
  min_words(mn);
  // the period: the smallest divisor dv of L with rotl(t, dv) = t
  // (descending, the last match wins)
  int d[B];
#pragma unroll
  for (int b = 0; b < B; ++b) d[b] = L;
#pragma unroll 1
  for (int i = a.n_div - 1; i >= 0; --i) {
    const uint32_t sh = uint32_t(32 - a.divisors[i]);
#pragma unroll
    for (int b = 0; b < B; ++b)
      if ((__builtin_amdgcn_alignbit(hi[b], lo[b], sh) >> lowb) == t[b]) d[b] = a.divisors[i];
  }
  if (SYM == 1) {
#pragma unroll
    for (int b = 0; b < B; ++b) {
      rep[b] = mn[b] >> lowb;
      orbit[b] = d[b];
      const int j = SHIFT ? int(mn[b] & low) : 0;
      ci[b] = j ? L - j : 0;
      sign[b] = 1;
    }
    return;
  }
  uint32_t mnp[B];
#pragma unroll
  for (int b = 0; b < B; ++b) doubled(b, reflect_l(t[b], L));
  min_words(mnp);
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const uint32_t rep_o = mn[b] >> lowb, rep_p = mnp[b] >> lowb;
    const bool p = rep_p < rep_o;
    rep[b] = p ? rep_p : rep_o;
    orbit[b] = rep_p == rep_o ? d[b] : 2 * d[b];
    const int j = SHIFT ? int((p ? mnp[b] : mn[b]) & low) : 0;
    ci[b] = j ? L - j : 0;
    sign[b] = p ? a.parity : 1;
  }
}

// General symmetry group: rep = min (or max) of g(t) over the elements, the
// orbit size |G| / |Stab(t)| and the phase χ(h⁻¹) of the first element h
// with h(t) = rep (t = h⁻¹ rep). Elements whose images tie differ by a
// stabilizer element, whose character is 1 on every orbit of the sector.
template <typename ST>
__device__ __forceinline__ void general_walk(ST t, const HvArgs& a, ST& rep, int& orbit, int& ci) {
  rep = t;
  int best = 0, stab = 0;
  // a rotation/reflection group of the ring: each element is the state, or
  // its bit reversal, rotated left by k (then the digit flips)
  const int W = 8 * int(sizeof(ST));
  ST tr = 0;
  if (a.grot) {
    if constexpr (sizeof(ST) == 4) tr = __brev(t) >> (W - a.L);
    else tr = __brevll(t) >> (W - a.L);
  }
  const ST mask = ST(a.mask64);
  for (int g = 0; g < a.ngroup; ++g) {
    ST u;
    if (a.grot) {
      const uint8_t code = a.grot[g];
      const ST v = (code & 0x40) ? tr : t;
      const int k = code & 63;
      u = k ? ST(((v << k) | (v >> (a.L - k))) & mask) : v;
    } else {
      const uint8_t* d = a.gdest + 64 * g;
      u = 0;
      for (int i = 0; i < a.L; ++i) u |= ((t >> i) & ST(1)) << d[i];
    }
    u ^= ST(a.gflip[g]);
    stab += u == t ? 1 : 0;
    if (a.maxrep ? u > rep : u < rep) {
      rep = u;
      best = g;
    }
  }
  orbit = a.ngroup / max(stab, 1);
  ci = a.gphase[best];
}

template <typename S>
struct Cx {
  S re, im;
};

// CMODE 0: real vectors, real coefficients
// CMODE 1: complex vectors, real coefficients (real characters)
// CMODE 2: complex vectors, complex coefficients/characters
template <typename S, int CMODE, int SYM, bool SHIFT, int NV, bool IMG>
__global__ void __launch_bounds__(kBlock) hv_kernel(HvArgs a, const S* __restrict__ x, S* __restrict__ y) {
  if (a.gate && *a.gate) return;
  extern __shared__ double smem_d[];
  double* s_amp = smem_d;                                  // n_amp or 2*n_amp
  double* s_chr = s_amp + (CMODE == 2 ? 2 : 1) * a.n_amp;  // 2*nchar
  double* s_sq = s_chr + 2 * kMaxChars;                    // 65
  double* s_rsq = s_sq + 65;                               // 65
  double* s_crsq = s_rsq + 65;                             // 16 (image mode)
  uint32_t* s_meta = reinterpret_cast<uint32_t*>(s_crsq + 16);
  uint32_t* s_hop = s_meta + 4 * (a.n_diag + a.n_off);     // G (image mode)
  uint32_t* s_bin = s_hop + (a.n_diag + a.n_off);          // 1024 (image mode)
  uint32_t* s_cells = IMG ? s_bin + 1024 : s_hop;          // 2*ncells (walk mode; 8-byte aligned)

  const int G = a.n_diag + a.n_off;
  const int namp = (CMODE == 2 ? 2 : 1) * a.n_amp;
  for (int i = threadIdx.x; i < namp; i += kBlock) s_amp[i] = a.amp[i];
  for (int i = threadIdx.x; i < 4 * G; i += kBlock) s_meta[i] = a.meta[i];
  if (IMG) {
    for (int i = threadIdx.x; i < G; i += kBlock) s_hop[i] = a.hop[i];
    if (a.up >= 0)
      for (int i = threadIdx.x; i < 1024; i += kBlock) s_bin[i] = a.binom[i];
    for (int i = threadIdx.x; i < 16; i += kBlock) s_crsq[i] = a.code_rsq[i];
  } else {
    for (int i = threadIdx.x; i < 2 * a.ncells; i += kBlock) s_cells[i] = a.dcells[i];
  }
  if (SYM == 1 || SYM == 3)
    for (int i = threadIdx.x; i < 2 * a.nchar; i += kBlock) s_chr[i] = a.chars[i];
  for (int i = threadIdx.x; i < 65; i += kBlock) {
    s_sq[i] = sqrt(double(i));
    s_rsq[i] = i ? 1.0 / sqrt(double(i)) : 0.0;
  }
  __syncthreads();

  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t row = gid / uint32_t(a.lpr);
  const int sub = int(gid % uint32_t(a.lpr));
  if (row >= a.n) return;  // all lanes of a row leave together
  const uint32_t r = a.rstates[row];
  S sq_r = 1;
  if (SYM) sq_r = S(s_sq[a.orbits[row]]);
  uint32_t rank_r = 0;
  if (IMG && a.up >= 0) rank_r = rank_colex(r, s_bin);

  // diagonal (first lane of the row): the per-row table of the current
  // coefficients when there is one, else a walk over the diagonal groups
  Cx<S> dsum{0, 0};
  if (a.diag) {
    if (sub == 0) {
      dsum.re = S(a.diag[row]);
      if (CMODE == 2) dsum.im = S(a.diag[a.n + row]);
    }
  } else {
    const int ndiag = sub == 0 ? a.n_diag : 0;
    for (int g = 0; g < ndiag; ++g) {
      const int pat = pattern_of(r, s_meta[4 * g + 1]);
      const int off = int(s_meta[4 * g + 3]) + pat;
      if (CMODE == 2) {
        dsum.re += S(s_amp[2 * off]);
        dsum.im += S(s_amp[2 * off + 1]);
      } else {
        dsum.re += S(s_amp[off]);
      }
    }
  }
  Cx<S> acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (CMODE == 0) {
      acc[v].re = dsum.re * x[v * a.ld + row];
      acc[v].im = 0;
    } else {
      const S xr = x[2 * (v * a.ld + row)], xi = x[2 * (v * a.ld + row) + 1];
      acc[v].re = dsum.re * xr - dsum.im * xi;
      acc[v].im = dsum.re * xi + dsum.im * xr;
    }
  }

  // off-diagonal
  if constexpr (IMG) {
    // Batched: the image words of up to B targets are loaded together, then
    // the B input entries they name, then the B products are accumulated —
    // two memory round trips per batch instead of two per target.
    constexpr int B = NV > 1 ? 4 : 8;
    const uint32_t imask = (1u << a.img_bits) - 1u;
    for (int w0 = 0; w0 < a.n_off; w0 += 32) {
      uint32_t act = 0;
      const int gend = min(32, a.n_off - w0);
      for (int gi = 0; gi < gend; ++gi) {
        const int g = a.n_diag + w0 + gi;
        const int pat = pattern_of(r, s_meta[4 * g + 1]);
        act |= ((s_meta[4 * g + 2] >> pat) & 1u) << gi;
      }
      act &= a.lane_groups << sub;
      while (act) {
        // ranks of up to B targets (ALU and LDS only), then their image words
        uint32_t ent[B], rts[B];
        int offs[B];
        bool in[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
          in[b] = false;
          rts[b] = 0;
          offs[b] = 0;
          if (act) {
            const int gi = __ffs(act) - 1;
            act &= act - 1;
            const int g = a.n_diag + w0 + gi;
            const uint32_t t = r ^ s_meta[4 * g];
            offs[b] = int(s_meta[4 * g + 3]) + pattern_of(r, s_meta[4 * g + 1]);
            uint32_t rt = t;
            bool ok = true;
            if (a.up >= 0) {
              const uint32_t hp = s_hop[g];
              if (hp >> 16) {
                const int p = hp & 255, q = (hp >> 8) & 255;
                rt = ((r >> p) & 1u) ? rank_r + uint32_t(move_delta(r, p, q, s_bin))
                                     : rank_r - uint32_t(move_delta(t, p, q, s_bin));
              } else {
                ok = __popc(t) == a.up;  // else it leaves the fixed-magnetization sector
                rt = ok ? rank_colex(t, s_bin) : 0u;
              }
            }
            in[b] = ok;
            rts[b] = rt;
          }
        }
#pragma unroll
        for (int b = 0; b < B; ++b) ent[b] = in[b] ? __ldg(a.img + rts[b]) : 0xFFFFFFFFu;
        S xr[B][NV], xi[B][NV];
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const int idx = ent[b] != 0xFFFFFFFFu ? int(ent[b] & imask) : int(row);
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            if (CMODE == 0) {
              xr[b][v] = x[v * a.ld + idx];
              xi[b][v] = 0;
            } else {
              xr[b][v] = x[2 * (v * a.ld + idx)];
              xi[b][v] = x[2 * (v * a.ld + idx) + 1];
            }
          }
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
          if (ent[b] == 0xFFFFFFFFu) continue;
          const uint32_t e = ent[b];
          const int code = int((e >> a.img_bits) & 15u);
          const int ci = int((e >> (a.img_bits + 5)) & 31u);
          S scale = ((e >> (a.img_bits + 4)) & 1u) ? S(-1) : S(1);
          if (SYM) scale *= sq_r * S(s_crsq[code]);
          Cx<S> c;
          if (CMODE == 2) {
            c.re = S(s_amp[2 * offs[b]]) * scale;
            c.im = S(s_amp[2 * offs[b] + 1]) * scale;
          } else {
            c.re = S(s_amp[offs[b]]) * scale;
            c.im = 0;
          }
          if ((SYM == 1 || SYM == 3) && SHIFT) {
            const S pr = S(s_chr[2 * ci]), pi = S(s_chr[2 * ci + 1]);
            if (CMODE == 2) {
              const S re = c.re * pr - c.im * pi;
              c.im = c.re * pi + c.im * pr;
              c.re = re;
            } else {
              c.re *= pr;  // real characters (k = L/2): pi == 0
            }
          }
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            if (CMODE == 0) {
              acc[v].re += c.re * xr[b][v];
            } else {
              acc[v].re += c.re * xr[b][v] - c.im * xi[b][v];
              acc[v].im += c.re * xi[b][v] + c.im * xr[b][v];
            }
          }
        }
      }
    }
  } else {
  const int nmax = a.max_sites;
  for (int w0 = 0; w0 < a.n_off; w0 += 32) {
    uint32_t act = 0;
    const int gend = min(32, a.n_off - w0);
    // g is uniform here: the group words go to scalar registers
    for (int gi = 0; gi < gend; ++gi) {
      const int g = a.n_diag + w0 + gi;
      const uint32_t packed = __builtin_amdgcn_readfirstlane(s_meta[4 * g + 1]);
      const uint32_t nz = __builtin_amdgcn_readfirstlane(s_meta[4 * g + 2]);
      const int pat = pattern_upto(r, packed, nmax);
      act |= ((nz >> pat) & 1u) << gi;
    }
    act &= a.lane_groups << sub;
    // Batched as in the image mode: the walks of up to B targets (vector
    // ALU, rotation loops shared), then their directory slots, then the B
    // input entries, then the B products in target order (the order of the
    // one-at-a-time loop, so the sums are the same). Four targets keep the
    // kernel under 64 registers (eight waves per SIMD), which measured
    // faster than eight targets at five waves.
    constexpr int B = 4;
    while (act) {
      uint32_t tgt[B], key[B];
      bool in[B];
      int offs[B], cis[B], sgn[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        in[b] = false;
        tgt[b] = r;
        offs[b] = 0;
        if (act) {
          const int gi = __ffs(act) - 1;
          act &= act - 1;
          const int g = a.n_diag + w0 + gi;
          tgt[b] = r ^ s_meta[4 * g];
          offs[b] = int(s_meta[4 * g + 3]) + pattern_upto(r, s_meta[4 * g + 1], nmax);
          in[b] = true;
        }
      }
      int orb[B];
      if constexpr (SYM != 0) {
        walk_batch<SYM, SHIFT, B>(tgt, a, key, orb, cis, sgn);
      } else {
#pragma unroll
        for (int b = 0; b < B; ++b) {
          key[b] = tgt[b];
          orb[b] = 1;
          cis[b] = 0;
          sgn[b] = 1;
        }
      }
      int idx[B];
      dir_find<B>(key, in, s_cells, a, idx);
      S xr[B][NV], xi[B][NV];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        const int at = idx[b] >= 0 ? idx[b] : int(row);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          if (CMODE == 0) {
            xr[b][v] = x[v * a.ld + at];
            xi[b][v] = 0;
          } else {
            xr[b][v] = x[2 * (v * a.ld + at)];
            xi[b][v] = x[2 * (v * a.ld + at) + 1];
          }
        }
      }
#pragma unroll
      for (int b = 0; b < B; ++b) {
        if (idx[b] < 0) continue;
        Cx<S> c;
        if (CMODE == 2) {
          c.re = S(s_amp[2 * offs[b]]);
          c.im = S(s_amp[2 * offs[b] + 1]);
        } else {
          c.re = S(s_amp[offs[b]]);
          c.im = 0;
        }
        S scale = S(sgn[b]);
        if (SYM) scale *= sq_r * S(s_rsq[orb[b]]);
        c.re *= scale;
        c.im *= scale;
        if ((SYM == 1 || SYM == 3) && SHIFT) {
          const S pr = S(s_chr[2 * cis[b]]), pi = S(s_chr[2 * cis[b] + 1]);
          if (CMODE == 2) {
            const S re = c.re * pr - c.im * pi;
            c.im = c.re * pi + c.im * pr;
            c.re = re;
          } else {
            c.re *= pr;  // real characters (k = L/2): pi == 0
          }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          if (CMODE == 0) {
            acc[v].re += c.re * xr[b][v];
          } else {
            acc[v].re += c.re * xr[b][v] - c.im * xi[b][v];
            acc[v].im += c.re * xi[b][v] + c.im * xr[b][v];
          }
        }
      }
    }
  }
  }
  for (int o = 1; o < a.lpr; o <<= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      acc[v].re += __shfl_xor(acc[v].re, o, 64);
      if (CMODE != 0) acc[v].im += __shfl_xor(acc[v].im, o, 64);
    }
  }
  if (sub != 0) return;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (CMODE == 0) {
      y[v * a.ld + row] = acc[v].re;
    } else {
      y[2 * (v * a.ld + row)] = acc[v].re;
      y[2 * (v * a.ld + row) + 1] = acc[v].im;
    }
  }
}

// Exact column entries of the sector matrix: lane r applies every coupling
// to its representative and merges the contributions per target row, which
// is the reference's column-streaming assembly for column r
// (src/operator/mod.rs:2119-2170). `amp` holds the forward amplitudes c·a as
// (re, im) pairs; the character enters conjugated (conj(phase_t)).
template <int CAP>
struct ColList {
  int rows[CAP];
  double vre[CAP], vim[CAP];
  int cnt = 0;
  double dre = 0, dim = 0;
  bool overflow = false;
};

// Wide sectors (64-bit states): the translation/reflection orbit walk one
// rotation at a time (rep, orbit size, phase = sign·χ(ci) as in walk_scan),
// and a plain binary search of the ascending representatives; these serve
// the one-time build kernels only (products run on the compiled stream).
__device__ __forceinline__ uint64_t rot64(uint64_t s, int j, int L, uint64_t mask) {
  return j ? (((s << j) | (s >> (L - j))) & mask) : s;
}

__device__ void walk64(uint64_t t, const HvArgs& a, int sym, uint64_t& rep, int& orbit, int& ci, int& sign) {
  const int L = a.L;
  sign = 1;
  ci = 0;
  const uint64_t ps = __brevll(t) >> (64 - L);
  if (sym == 2) {
    if (ps < t) {
      rep = ps;
      orbit = 2;
      sign = a.parity;
    } else {
      rep = t;
      orbit = (ps == t) ? 1 : 2;
    }
    return;
  }
  uint64_t mn = t;
  int jmin = 0, d = L;
  for (int j = 1; j < L; ++j) {
    const uint64_t r = rot64(t, j, L, a.mask64);
    if (r < mn) {
      mn = r;
      jmin = j;
    }
    if (r == t && d == L) d = j;
  }
  if (sym == 1) {
    rep = mn;
    orbit = d;
    ci = (L - jmin) % L;
    return;
  }
  uint64_t mnp = ps;
  int jminp = 0;
  for (int j = 1; j < L; ++j) {
    const uint64_t r = rot64(ps, j, L, a.mask64);
    if (r < mnp) {
      mnp = r;
      jminp = j;
    }
  }
  if (mnp < mn) {
    rep = mnp;
    orbit = 2 * d;
    ci = (L - jminp) % L;
    sign = a.parity;
  } else {
    rep = mn;
    orbit = (mnp == mn) ? d : 2 * d;
    ci = (L - jmin) % L;
  }
}

__device__ __forceinline__ int find_state64(uint64_t key, const uint64_t* __restrict__ states, uint32_t n) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (states[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && states[lo] == key) ? int(lo) : -1;
}

// the accessors of a row's state and of a group's flip mask, site pattern,
// nonzero-pattern mask and amplitude offset, for either state width
template <typename ST>
__device__ __forceinline__ ST row_state(const HvArgs& a, uint32_t row) {
  if constexpr (sizeof(ST) == 4) return a.rstates[row];
  else return a.rstates64[row];
}

__device__ __forceinline__ int pattern_of64(uint64_t r, uint64_t packed) {
  const int ns = int(packed & 7u);
  int pat = 0;
  for (int i = 0; i < ns; ++i) {
    const int site = int((packed >> (3 + 6 * i)) & 63u);
    pat |= int((r >> site) & 1u) << i;
  }
  return pat;
}

template <typename ST>
__device__ __forceinline__ ST g_flip(const HvArgs& a, const uint32_t* meta, int g) {
  if constexpr (sizeof(ST) == 4) return meta[4 * g];
  else return a.meta64[4 * g];
}
template <typename ST>
__device__ __forceinline__ int g_pattern(const HvArgs& a, const uint32_t* meta, ST r, int g) {
  if constexpr (sizeof(ST) == 4) return pattern_of(r, meta[4 * g + 1]);
  else return pattern_of64(r, a.meta64[4 * g + 1]);
}
template <typename ST>
__device__ __forceinline__ uint32_t g_nz(const HvArgs& a, const uint32_t* meta, int g) {
  if constexpr (sizeof(ST) == 4) return meta[4 * g + 2];
  else return uint32_t(a.meta64[4 * g + 2]);
}
template <typename ST>
__device__ __forceinline__ int g_off(const HvArgs& a, const uint32_t* meta, int g) {
  if constexpr (sizeof(ST) == 4) return int(meta[4 * g + 3]);
  else return int(a.meta64[4 * g + 3]);
}

template <typename ST>
__device__ __forceinline__ bool target_of(const HvArgs& a, int sym, int shift_needed, ST t, int& idx, int& orbit,
                                          int& ci, int& sign) {
  if constexpr (sizeof(ST) == 8) {
    uint64_t rep = t;
    orbit = 1;
    ci = 0;
    sign = 1;
    if (sym == 4) general_walk(t, a, rep, orbit, ci);
    else if (sym) walk64(t, a, sym, rep, orbit, ci, sign);
    idx = find_state64(rep, a.states64, a.n);
    if (idx >= 0 && a.rev) idx = int(a.n) - 1 - idx;
    return idx >= 0;
  } else {
  uint32_t rep = t;
  orbit = 1;
  ci = 0;
  sign = 1;
  if (sym == 1) {
    if (shift_needed) walk<1, true>(t, a, rep, orbit, ci, sign);
    else walk<1, false>(t, a, rep, orbit, ci, sign);
  } else if (sym == 2) {
    walk<2, false>(t, a, rep, orbit, ci, sign);
  } else if (sym == 3) {
    if (shift_needed) walk<3, true>(t, a, rep, orbit, ci, sign);
    else walk<3, false>(t, a, rep, orbit, ci, sign);
  } else if (sym == 4) {
    general_walk(t, a, rep, orbit, ci);
  }
  idx = dir_find1(rep, a);
  if (idx >= 0 && a.rev) idx = int(a.n) - 1 - idx;
  return idx >= 0;
  }
}

// contribution value c·a·sqrt(N_src/N_t)·conj(sign·χ(ci))
__device__ __forceinline__ void column_value(const HvArgs& a, const double* famp, int off, int sym, double nsrc,
                                             int orbit, int ci, int sign, double& re, double& im) {
  double cre = famp[2 * off], cim = famp[2 * off + 1];
  const double scale = double(sign) * (sym ? sqrt(nsrc / double(orbit)) : 1.0);
  cre *= scale;
  cim *= scale;
  if (sym == 1 || sym == 3 || sym == 4) {
    const double pr = a.chars[2 * ci], pi = -a.chars[2 * ci + 1];
    re = cre * pr - cim * pi;
    im = cre * pi + cim * pr;
  } else {
    re = cre;
    im = cim;
  }
}

template <typename ST, int CAP>
__device__ void collect_column(const HvArgs& a, const uint32_t* meta, const double* famp, int sym, int shift_needed,
                               uint32_t col, ColList<CAP>& out) {
  const ST r = row_state<ST>(a, col);
  const double ns = sym ? double(a.orbits[col]) : 1.0;
  for (int g = 0; g < a.n_diag; ++g) {
    const int off = g_off<ST>(a, meta, g) + g_pattern<ST>(a, meta, r, g);
    out.dre += famp[2 * off];
    out.dim += famp[2 * off + 1];
  }
  const int G = a.n_diag + a.n_off;
  for (int g = a.n_diag; g < G; ++g) {
    const int pat = g_pattern<ST>(a, meta, r, g);
    if (!((g_nz<ST>(a, meta, g) >> pat) & 1u)) continue;
    int idx, orbit, ci, sign;
    if (!target_of<ST>(a, sym, shift_needed, ST(r ^ g_flip<ST>(a, meta, g)), idx, orbit, ci, sign)) continue;
    double re, im;
    column_value(a, famp, g_off<ST>(a, meta, g) + pat, sym, ns, orbit, ci, sign, re, im);
    if (uint32_t(idx) == col) {
      out.dre += re;
      out.dim += im;
      continue;
    }
    int k = 0;
    while (k < out.cnt && out.rows[k] != idx) ++k;
    if (k == out.cnt) {
      if (out.cnt == CAP) {
        out.overflow = true;
        return;
      }
      out.rows[out.cnt] = idx;
      out.vre[out.cnt] = 0;
      out.vim[out.cnt] = 0;
      ++out.cnt;
    }
    out.vre[k] += re;
    out.vim[k] += im;
  }
}

// merged entry H[row, col] (all contributions of column col landing on row)
template <typename ST>
__device__ void entry_at(const HvArgs& a, const uint32_t* meta, const double* famp, int sym, int shift_needed,
                         uint32_t col, uint32_t row, double& vre, double& vim) {
  const ST r = row_state<ST>(a, col);
  const double ns = sym ? double(a.orbits[col]) : 1.0;
  vre = 0;
  vim = 0;
  if (row == col) {
    for (int g = 0; g < a.n_diag; ++g) {
      const int off = g_off<ST>(a, meta, g) + g_pattern<ST>(a, meta, r, g);
      vre += famp[2 * off];
      vim += famp[2 * off + 1];
    }
  }
  const int G = a.n_diag + a.n_off;
  for (int g = a.n_diag; g < G; ++g) {
    const int pat = g_pattern<ST>(a, meta, r, g);
    if (!((g_nz<ST>(a, meta, g) >> pat) & 1u)) continue;
    int idx, orbit, ci, sign;
    if (!target_of<ST>(a, sym, shift_needed, ST(r ^ g_flip<ST>(a, meta, g)), idx, orbit, ci, sign)) continue;
    if (uint32_t(idx) != row) continue;
    double re, im;
    column_value(a, famp, g_off<ST>(a, meta, g) + pat, sym, ns, orbit, ci, sign, re, im);
    vre += re;
    vim += im;
  }
}

constexpr int kMaxColEntries = 128;
constexpr double kEps = 2.220446049250313e-16;

// dense column write / entry counts / diagonal / absolute column sums
template <typename ST>
__global__ void __launch_bounds__(kBlock) column_entries_kernel(HvArgs a, const double* famp, int sym, int shift_needed,
                                                                cplx* dense, double* colsum, double* diag_re,
                                                                double* diag_im, int* overflow, cplx shift,
                                                                uint32_t* count) {
  const uint32_t col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= a.n) return;
  ColList<kMaxColEntries> C;
  collect_column<ST>(a, a.meta, famp, sym, shift_needed, col, C);
  if (C.overflow) {
    *overflow = 1;
    return;
  }
  if (count) {
    // stored entries of the reference matrix: merged values with |v| > eps
    uint32_t c = (hypot(C.dre, C.dim) > kEps) ? 1u : 0u;
    for (int k = 0; k < C.cnt; ++k)
      if (hypot(C.vre[k], C.vim[k]) > kEps) ++c;
    count[col] = c;
    return;
  }
  if (dense) {  // row-major: dense[row * n + col]
    dense[size_t(col) * a.n + col] = cplx(C.dre, C.dim);
    for (int k = 0; k < C.cnt; ++k) dense[size_t(C.rows[k]) * a.n + col] = cplx(C.vre[k], C.vim[k]);
    return;
  }
  if (diag_re) {
    diag_re[col] = C.dre;
    diag_im[col] = C.dim;
  }
  if (colsum) {
    // shifted_trace_and_one_norm (src/solve/mod.rs:59-104): off-diagonal
    // stored entries |v| plus |diag - shift| for every column
    double s = 0;
    for (int k = 0; k < C.cnt; ++k) {
      double m = hypot(C.vre[k], C.vim[k]);
      if (m > kEps) s += m;
    }
    double dr = (hypot(C.dre, C.dim) > kEps) ? C.dre : 0.0, di = (hypot(C.dre, C.dim) > kEps) ? C.dim : 0.0;
    s += hypot(dr - shift.real(), di - shift.imag());
    colsum[col] = s;
  }
}

// Stored sector matrix (the reference's assembled formats,
// src/operator/mod.rs:2091-2216): column `col`'s stored entries (merged per
// row, |v| > eps, the diagonal included) sorted by row at ptr[col]: the CSC
// arrays. The count pass is column_entries_kernel's count mode, which keeps
// exactly the same entries.
template <typename ST>
__global__ void __launch_bounds__(kBlock) column_fill_kernel(HvArgs a, const double* famp, int sym, int shift_needed,
                                                             const uint64_t* __restrict__ ptr, uint32_t* rows,
                                                             double2* vals, int* overflow) {
  const uint32_t col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= a.n) return;
  ColList<kMaxColEntries> C;
  collect_column<ST>(a, a.meta, famp, sym, shift_needed, col, C);
  if (C.overflow) {
    *overflow = 1;
    return;
  }
  // keep the stored off-diagonal entries, then order them by row
  int m = 0;
  for (int k = 0; k < C.cnt; ++k) {
    if (hypot(C.vre[k], C.vim[k]) <= kEps) continue;
    C.rows[m] = C.rows[k];
    C.vre[m] = C.vre[k];
    C.vim[m] = C.vim[k];
    ++m;
  }
  for (int k = 1; k < m; ++k) {
    const int r = C.rows[k];
    const double vr = C.vre[k], vi = C.vim[k];
    int j = k - 1;
    while (j >= 0 && C.rows[j] > r) {
      C.rows[j + 1] = C.rows[j];
      C.vre[j + 1] = C.vre[j];
      C.vim[j + 1] = C.vim[j];
      --j;
    }
    C.rows[j + 1] = r;
    C.vre[j + 1] = vr;
    C.vim[j + 1] = vi;
  }
  const bool diag = hypot(C.dre, C.dim) > kEps;
  uint64_t out = ptr[col];
  bool placed = !diag;
  for (int k = 0; k < m; ++k) {
    if (!placed && uint32_t(C.rows[k]) > col) {
      rows[out] = col;
      vals[out++] = make_double2(C.dre, C.dim);
      placed = true;
    }
    rows[out] = uint32_t(C.rows[k]);
    vals[out++] = make_double2(C.vre[k], C.vim[k]);
  }
  if (!placed) {
    rows[out] = col;
    vals[out] = make_double2(C.dre, C.dim);
  }
}

// the column of every CSC entry
__global__ void csc_columns_kernel(const uint64_t* __restrict__ ptr, uint32_t n, uint32_t* cols) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (uint64_t k = ptr[c]; k < ptr[c + 1]; ++k) cols[k] = c;
}

// CSR values from CSC entries: Hermitian matrices read column r as the
// conjugate of row r (perm == nullptr: same positions); otherwise entry k of
// the CSR is CSC entry perm[k] (stable order by row, then column)
__global__ void csr_values_kernel(const double2* __restrict__ vals, const uint32_t* __restrict__ perm,
                                  const uint32_t* __restrict__ ccols, uint64_t nnz, int complex_out, double* out,
                                  uint32_t* cols) {
  const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint64_t src = perm ? perm[k] : k;
  const double2 v = vals[src];
  if (perm) cols[k] = ccols[src];
  if (complex_out) {
    out[2 * k] = v.x;
    out[2 * k + 1] = perm ? v.y : -v.y;
  } else {
    out[k] = v.x;
  }
}

// row counts of sorted row keys
__global__ void row_count_kernel(const uint32_t* __restrict__ rows, uint64_t nnz, uint32_t* counts) {
  const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < nnz) atomicAdd(counts + rows[k], 1u);
}

// Hermiticity from the stored entries: entry k of H in row-major order is
// (keys[k], ccols[perm[k]], vals[perm[k]]); CSC entry k read as H^H is
// (ccols[k], rows[k], conj(vals[k])) in the same order. Equal positions
// and |v - conj(w)| <= 1e-12 everywhere make H Hermitian (flags[0] stays
// 0); a position mismatch (an entry the other side lacks) sets flags[1].
__global__ void herm_compare_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ perm,
                                    const uint32_t* __restrict__ ccols, const uint32_t* __restrict__ rows,
                                    const double2* __restrict__ vals, uint64_t nnz, int* flags) {
  const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint32_t src = perm[k];
  if (keys[k] != ccols[k] || ccols[src] != rows[k]) {
    flags[1] = 1;
    return;
  }
  const double2 v = vals[src], w = vals[k];
  if (hypot(v.x - w.x, v.y + w.y) > 1.0e-12) flags[0] = 1;
}

// y = A x from the stored CSR: 8 lanes per row, each summing every 8th
// entry of the row, then a fixed butterfly over the 8 lanes (deterministic).
// CMODE 0: real values and vectors; 1: real values, complex vectors; 2:
// complex values and vectors. SX: input type, S: accumulation and output.
constexpr int kCsrLanes = 8;

template <typename S, int CMODE, typename SX>
__global__ void __launch_bounds__(kBlock) csr_kernel(const uint64_t* __restrict__ ptr,
                                                     const uint32_t* __restrict__ col,
                                                     const double* __restrict__ val, uint32_t n,
                                                     const SX* __restrict__ x, S* __restrict__ y,
                                                     const int* gate) {
  if (gate && *gate) return;
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t row = gid / kCsrLanes;
  const int sub = int(gid % kCsrLanes);
  S re = 0, im = 0;
  if (row < n) {
    const uint64_t end = ptr[row + 1];
    for (uint64_t k = ptr[row] + uint64_t(sub); k < end; k += kCsrLanes) {
      const uint32_t c = col[k];
      if (CMODE == 0) {
        re += S(val[k]) * S(x[c]);
      } else if (CMODE == 1) {
        const S v = S(val[k]);
        re += v * S(x[2 * size_t(c)]);
        im += v * S(x[2 * size_t(c) + 1]);
      } else {
        const S vr = S(val[2 * k]), vi = S(val[2 * k + 1]);
        const S xr = S(x[2 * size_t(c)]), xi = S(x[2 * size_t(c) + 1]);
        re += vr * xr - vi * xi;
        im += vr * xi + vi * xr;
      }
    }
  }
#pragma unroll
  for (int o = kCsrLanes / 2; o > 0; o >>= 1) {
    re += __shfl_xor(re, o, kCsrLanes);
    if (CMODE != 0) im += __shfl_xor(im, o, kCsrLanes);
  }
  if (row >= n || sub != 0) return;
  if (CMODE == 0) {
    y[row] = re;
  } else {
    y[2 * size_t(row)] = re;
    y[2 * size_t(row) + 1] = im;
  }
}

// Hermiticity of the sector matrix, entry by entry as the reference checks
// it: every stored entry (j, r, v) needs |v - conj(H[r, j])| <= 1e-12
// (src/operator/mod.rs:1240-1242, 1313-1398, 2212-2214). Lane r lists column
// r and re-evaluates column j for each of its entries.
template <typename ST>
__global__ void __launch_bounds__(kBlock) hermiticity_kernel(HvArgs a, const double* famp, int sym, int shift_needed,
                                                             int* flags) {
  const uint32_t col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= a.n) return;
  ColList<kMaxColEntries> C;
  collect_column<ST>(a, a.meta, famp, sym, shift_needed, col, C);
  if (C.overflow) {
    flags[1] = 1;
    return;
  }
  const double tol = 1.0e-12;
  bool bad = false;
  if (hypot(C.dre, C.dim) > kEps && 2.0 * fabs(C.dim) > tol) bad = true;
  for (int k = 0; k < C.cnt && !bad; ++k) {
    if (hypot(C.vre[k], C.vim[k]) <= kEps) continue;  // not stored
    double wr, wi;
    entry_at<ST>(a, a.meta, famp, sym, shift_needed, uint32_t(C.rows[k]), col, wr, wi);
    if (hypot(wr, wi) <= kEps) wr = wi = 0.0;
    if (hypot(C.vre[k] - wr, C.vim[k] + wi) > tol) bad = true;
  }
  if (bad) flags[0] = 1;
}

// y = H x by columns with atomic accumulation, for sector matrices that are
// not Hermitian (assembly checks disabled); y must be zeroed first.
template <typename ST>
__global__ void __launch_bounds__(kBlock) scatter_kernel(HvArgs a, const double* famp, int sym, int shift_needed,
                                                         const double2* __restrict__ x, double* y) {
  const uint32_t col = blockIdx.x * kBlock + threadIdx.x;
  if (col >= a.n) return;
  const ST r = row_state<ST>(a, col);
  const double ns = sym ? double(a.orbits[col]) : 1.0;
  const double2 xc = x[col];
  double dre = 0, dim_ = 0;
  for (int g = 0; g < a.n_diag; ++g) {
    const int off = g_off<ST>(a, a.meta, g) + g_pattern<ST>(a, a.meta, r, g);
    dre += famp[2 * off];
    dim_ += famp[2 * off + 1];
  }
  atomicAdd(y + 2 * col, dre * xc.x - dim_ * xc.y);
  atomicAdd(y + 2 * col + 1, dre * xc.y + dim_ * xc.x);
  const int G = a.n_diag + a.n_off;
  for (int g = a.n_diag; g < G; ++g) {
    const int pat = g_pattern<ST>(a, a.meta, r, g);
    if (!((g_nz<ST>(a, a.meta, g) >> pat) & 1u)) continue;
    int idx, orbit, ci, sign;
    if (!target_of<ST>(a, sym, shift_needed, ST(r ^ g_flip<ST>(a, a.meta, g)), idx, orbit, ci, sign)) continue;
    double re, im;
    column_value(a, famp, g_off<ST>(a, a.meta, g) + pat, sym, ns, orbit, ci, sign, re, im);
    atomicAdd(y + 2 * idx, re * xc.x - im * xc.y);
    atomicAdd(y + 2 * idx + 1, re * xc.y + im * xc.x);
  }
}

__device__ uint32_t unrank_colex32(uint32_t rank, int L, int up, const uint32_t* binom) {
  uint32_t s = 0;
  int hi = L - 1;
  for (int k = up; k >= 1; --k) {
    int lo = k - 1, h = hi;
    while (lo < h) {
      const int mid = (lo + h + 1) >> 1;
      if (binom[mid * 32 + k] <= rank) lo = mid;
      else h = mid - 1;
    }
    s |= 1u << lo;
    rank -= binom[lo * 32 + k];
    hi = lo - 1;
  }
  return s;
}

// Image table: for every parent state (by colex rank for a fixed up count,
// by value otherwise) the orbit walk + sorted-table lookup done once:
// idx | orbit code << bits | (sign < 0) << (bits + 4) | ci << (bits + 5).
constexpr int kImgChunk = 32;

__global__ void __launch_bounds__(kBlock) image_kernel(HvArgs a, int sym, int shift_needed, uint64_t popmask,
                                                       uint32_t n_parents, const uint8_t* orbit_code,
                                                       uint32_t* out) {
  const uint32_t first = (blockIdx.x * kBlock + threadIdx.x) * uint32_t(kImgChunk);
  if (first >= n_parents) return;
  const int todo = int(min<uint32_t>(kImgChunk, n_parents - first));
  uint32_t s = a.up >= 0 ? unrank_colex32(first, a.L, a.up, a.binom) : first;
  for (int i = 0; i < todo; ++i) {
    uint32_t e = 0xFFFFFFFFu;
    if (a.up >= 0 || ((popmask >> __popc(s)) & 1ull)) {
      int idx, orbit, ci, sign;
      if (target_of<uint32_t>(a, sym, shift_needed, s, idx, orbit, ci, sign))
        e = uint32_t(idx) | (uint32_t(orbit_code[orbit]) << a.img_bits) |
            (uint32_t(sign < 0) << (a.img_bits + 4)) | (uint32_t(ci) << (a.img_bits + 5));
    }
    out[first + i] = e;
    if (a.up >= 0) {
      if (i + 1 < todo) {
        const uint32_t low = s & (0u - s), ripple = s + low;
        s = (((ripple ^ s) >> 2) / low) | ripple;
      }
    } else {
      ++s;
    }
  }
}

// Compiled index stream: for every representative, the packed target word
// (as in the image table: idx | orbit code << bits | (sign < 0) << (bits+4) |
// ci << (bits+5), or 0xFFFFFFFF outside the sector) of each of its active
// off-diagonal flip groups, in group order. Sliced ELL over 64-row slices
// (one wavefront): slice s holds 64·width_s words and word k of its row l
// sits at slice_ptr[s] + 64·k + l, so the words of one ordinal are read by a
// wavefront as one contiguous 256-byte line; shorter rows are padded with
// 0xFFFFFFFF. The stream depends only on the sector and the flip-group
// structure, so an operator family keeps it across evaluations.
constexpr int kSlice = 64;

// the active off-diagonal groups of representative r in group order
// (ST: the state width; wide states read the 64-bit group words meta64)
template <typename ST>
struct ActiveGroupsT {
  const uint32_t* meta;
  const uint64_t* meta64;
  int n_diag, n_off;
  ST r;
  int w0;
  uint32_t act;
  __device__ ActiveGroupsT(const uint32_t* m, int nd, int no, ST rr, const uint64_t* m64 = nullptr)
      : meta(m), meta64(m64), n_diag(nd), n_off(no), r(rr), w0(-32), act(0) {}
  __device__ uint32_t window(int w) const {
    uint32_t out = 0;
    const int gend = min(32, n_off - w);
    for (int gi = 0; gi < gend; ++gi) {
      const int g = n_diag + w + gi;
      if constexpr (sizeof(ST) == 4) {
        const int pat = pattern_of(r, meta[4 * g + 1]);
        out |= ((meta[4 * g + 2] >> pat) & 1u) << gi;
      } else {
        const int pat = pattern_of64(r, meta64[4 * g + 1]);
        out |= uint32_t((meta64[4 * g + 2] >> pat) & 1u) << gi;
      }
    }
    return out;
  }
  // next active group, -1 when none is left
  __device__ int next() {
    while (!act) {
      w0 += 32;
      if (w0 >= n_off) return -1;
      act = window(w0);
    }
    const int gi = __ffs(act) - 1;
    act &= act - 1;
    return n_diag + w0 + gi;
  }
};
using ActiveGroups = ActiveGroupsT<uint32_t>;

// width_s = the largest active-group count among the slice's rows
template <typename ST>
__global__ void __launch_bounds__(kBlock) stream_count_kernel(HvArgs a, uint32_t n_slices, uint32_t* width) {
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  int cnt = 0;
  if (row < a.n) {
    ActiveGroupsT<ST> ag(a.meta, a.n_diag, a.n_off, row_state<ST>(a, row), a.meta64);
    for (int w = 0; w < a.n_off; w += 32) cnt += __popc(ag.window(w));
  }
  for (int o = 32; o > 0; o >>= 1) cnt = max(cnt, __shfl_xor(cnt, o, 64));
  if ((threadIdx.x & 63) == 0 && row / kSlice < n_slices) width[row / kSlice] = uint32_t(cnt);
}

template <typename ST>
__global__ void __launch_bounds__(kBlock) stream_fill_kernel(HvArgs a, int sym, int shift_needed, uint32_t n_slices,
                                                             const uint8_t* orbit_code, const uint32_t* slice_ptr,
                                                             uint32_t* out, uint8_t* aux_out) {
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t slice = row / kSlice, lane = row % kSlice;
  if (slice >= n_slices) return;
  const uint32_t base = slice_ptr[slice];
  const uint32_t width = (slice_ptr[slice + 1] - base) / kSlice;
  uint32_t k = 0;
  if (row < a.n) {
    const ST r = row_state<ST>(a, row);
    ActiveGroupsT<ST> ag(a.meta, a.n_diag, a.n_off, r, a.meta64);
    for (int g = ag.next(); g >= 0; g = ag.next()) {
      uint32_t e = 0xFFFFFFFFu;
      int idx, orbit, ci, sign;
      if (target_of<ST>(a, sym, shift_needed, ST(r ^ g_flip<ST>(a, a.meta, g)), idx, orbit, ci, sign)) {
        if (a.compact) {
          const uint32_t off = uint32_t(g_off<ST>(a, a.meta, g)) + uint32_t(g_pattern<ST>(a, a.meta, r, g)) -
                               uint32_t(a.c_off_base);
          // the fields above the entry, as the split layout's byte
          uint32_t hi = 0;
          int hs = 0;
          if (a.c_cb) hi |= uint32_t(orbit_code[orbit]) << hs;
          hs += a.c_cb;
          if (a.c_sb) hi |= uint32_t(sign < 0) << hs;
          hs += a.c_sb;
          if (shift_needed) hi |= uint32_t(ci) << hs;
          const int sc = a.c_b + a.c_ob;
          e = uint32_t(idx) | (off << a.c_b);
          if (aux_out) aux_out[base + kSlice * k + lane] = uint8_t(hi);
          else if (sc < 32) e |= hi << sc;
        } else {
          e = uint32_t(idx) | (uint32_t(orbit_code[orbit]) << a.img_bits) | (uint32_t(sign < 0) << (a.img_bits + 4)) |
              (uint32_t(ci) << (a.img_bits + 5));
        }
      }
      out[base + kSlice * k + lane] = e;
      ++k;
    }
  }
  for (; k < width; ++k) {
    out[base + kSlice * k + lane] = 0xFFFFFFFFu;
    if (aux_out) aux_out[base + kSlice * k + lane] = 0;
  }
}

// H·v from the compiled stream: one lane per row; the words of B ordinals
// are loaded together (coalesced over the wavefront), then the B input
// entries they name, then the B products are accumulated.
// The rows of block `blk` of one sector (shared by the single-sector kernel
// and the sector-batched kernel).
// S: accumulation and output type; SX: input type (float input with double
// accumulation and output is the mixed-precision product of the Lanczos
// FP32 schedule)
// IL: interleaved columns (element (row, v) at row · ld + v, the columns of
// a row contiguous, so the NV inputs a target names share cache lines);
// otherwise column-major (v · ld + row).
template <typename S, int CMODE, int SYM, bool SHIFT, int NV, typename SX = S, bool IL = false>
__device__ __forceinline__ void stream_rows(const HvArgs& a, const SX* __restrict__ x, S* __restrict__ y,
                                            uint32_t blk) {
  auto at = [&](int v, uint32_t i) -> size_t { return IL ? size_t(i) * a.ld + v : size_t(v) * a.ld + i; };
  extern __shared__ double smem_d[];
  double* s_amp = smem_d;                                  // n_amp or 2*n_amp
  double* s_chr = s_amp + (CMODE == 2 ? 2 : 1) * a.n_amp;  // 2*nchar
  double* s_sq = s_chr + 2 * kMaxChars;                    // 65
  double* s_crsq = s_sq + 65;                              // 16
  uint32_t* s_meta = reinterpret_cast<uint32_t*>(s_crsq + 16);
  const int G = a.n_diag + a.n_off;
  const int namp = (CMODE == 2 ? 2 : 1) * a.n_amp;
  for (int i = threadIdx.x; i < namp; i += kBlock) s_amp[i] = a.amp[i];
  for (int i = threadIdx.x; i < 4 * G; i += kBlock) s_meta[i] = a.meta[i];
  for (int i = threadIdx.x; i < 16; i += kBlock) s_crsq[i] = a.code_rsq[i];
  if (SYM == 1 || SYM == 3)
    for (int i = threadIdx.x; i < 2 * a.nchar; i += kBlock) s_chr[i] = a.chars[i];
  for (int i = threadIdx.x; i < 65; i += kBlock) s_sq[i] = sqrt(double(i));
  __syncthreads();

  const uint32_t row = blk * kBlock + threadIdx.x;
  const uint32_t slice = row / kSlice, lane = row % kSlice;
  const uint32_t n_slices = (a.n + kSlice - 1) / kSlice;
  if (slice >= n_slices) return;  // whole wavefronts leave together
  const bool live = row < a.n;
  const uint32_t base = a.slice_ptr[slice];
  const uint32_t width = (a.slice_ptr[slice + 1] - base) / kSlice;
  // the state itself is needed only to walk the groups
  const bool walk = !(a.compact && a.diag);
  const uint32_t r = (live && walk) ? a.rstates[row] : 0u;
  S sq_r = 1;
  if (SYM && live) {
    const int o = a.orbits[row];
    sq_r = S(o <= 64 ? s_sq[o] : sqrt(double(o)));  // general groups reach 255
  }

  Cx<S> acc[NV];
  {
    Cx<S> dsum{0, 0};
    if (a.diag) {
      if (live) {
        dsum.re = S(a.diag[row]);
        if (CMODE == 2) dsum.im = S(a.diag[a.n + row]);
      }
    } else {
      const int ndiag = live ? a.n_diag : 0;
      for (int g = 0; g < ndiag; ++g) {
        const int off = int(s_meta[4 * g + 3]) + pattern_of(r, s_meta[4 * g + 1]);
        if (CMODE == 2) {
          dsum.re += S(s_amp[2 * off]);
          dsum.im += S(s_amp[2 * off + 1]);
        } else {
          dsum.re += S(s_amp[off]);
        }
      }
    }
    const uint32_t own = live ? row : 0u;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (CMODE == 0) {
        acc[v].re = dsum.re * S(x[at(v, own)]);
        acc[v].im = 0;
      } else {
        const S xr = S(x[2 * at(v, own)]), xi = S(x[2 * at(v, own) + 1]);
        acc[v].re = dsum.re * xr - dsum.im * xi;
        acc[v].im = dsum.re * xi + dsum.im * xr;
      }
    }
  }
  constexpr int B = NV > 1 ? 4 : 8;
  const uint32_t* words = a.stream + base + lane;
  if (a.compact) {
    // every entry names its amplitude, orbit code, sign and character: no
    // group walk; loads are unconditional (clamped ordinal, then masked).
    // SPLIT: the fields above the amplitude entry come from the split
    // layout's byte per word instead of the word's top bits (two loop
    // bodies, so the 32-bit layout keeps its registers)
    const int so = a.c_b, sc = so + a.c_ob, ss = sc + a.c_cb, sh = ss + a.c_sb;
    const uint32_t imask = (1u << a.c_b) - 1u;
    auto run = [&](auto split_tag) {
      constexpr bool SPLIT = decltype(split_tag)::value;
      const uint8_t* auxw = SPLIT ? a.aux + base + lane : nullptr;
      for (uint32_t k0 = 0; k0 < width; k0 += B) {
        uint32_t ent[B];
        uint32_t hi[SPLIT ? B : 1];
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const size_t pos = size_t(kSlice) * min(k0 + b, width - 1);
          const uint32_t e = words[pos];
          ent[b] = k0 + b < width ? e : 0xFFFFFFFFu;
          if constexpr (SPLIT) hi[b] = uint32_t(auxw[pos]);
        }
        S xr[B][NV], xi[B][NV];
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const int idx = ent[b] != 0xFFFFFFFFu ? int(ent[b] & imask) : 0;
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            if (CMODE == 0) {
              xr[b][v] = S(x[at(v, idx)]);
              xi[b][v] = 0;
            } else {
              xr[b][v] = S(x[2 * at(v, idx)]);
              xi[b][v] = S(x[2 * at(v, idx) + 1]);
            }
          }
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const uint32_t e = ent[b];
          if (e == 0xFFFFFFFFu) continue;
          const int off = a.c_off_base + int(field(e, so, a.c_ob));
          uint32_t code, neg, ci = 0;
          if constexpr (SPLIT) {
            code = field(hi[b], 0, a.c_cb);
            neg = field(hi[b], a.c_cb, a.c_sb);
            if ((SYM == 1 || SYM == 3) && SHIFT) ci = hi[b] >> (a.c_cb + a.c_sb);
          } else {
            code = field(e, sc, a.c_cb);
            neg = field(e, ss, a.c_sb);
            if ((SYM == 1 || SYM == 3) && SHIFT) ci = e >> sh;
          }
          S scale = neg ? S(-1) : S(1);
          if (SYM) scale *= sq_r * S(s_crsq[code]);
          Cx<S> c;
          if (CMODE == 2) {
            c.re = S(s_amp[2 * off]) * scale;
            c.im = S(s_amp[2 * off + 1]) * scale;
          } else {
            c.re = S(s_amp[off]) * scale;
            c.im = 0;
          }
          if ((SYM == 1 || SYM == 3) && SHIFT) {
            const S pr = S(s_chr[2 * ci]), pi = S(s_chr[2 * ci + 1]);
            if (CMODE == 2) {
              const S re = c.re * pr - c.im * pi;
              c.im = c.re * pi + c.im * pr;
              c.re = re;
            } else {
              c.re *= pr;  // real characters (k = L/2): pi == 0
            }
          }
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            if (CMODE == 0) {
              acc[v].re += c.re * xr[b][v];
            } else {
              acc[v].re += c.re * xr[b][v] - c.im * xi[b][v];
              acc[v].im += c.re * xi[b][v] + c.im * xr[b][v];
            }
          }
        }
      }
    };
    if (a.aux) run(std::true_type{});
    else run(std::false_type{});
    if (!live) return;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (CMODE == 0) {
        y[at(v, row)] = acc[v].re;
      } else {
        y[2 * at(v, row)] = acc[v].re;
        y[2 * at(v, row) + 1] = acc[v].im;
      }
    }
    return;
  }
  const uint32_t imask = (1u << a.img_bits) - 1u;
  ActiveGroups ag(s_meta, a.n_diag, a.n_off, r);
  for (uint32_t k0 = 0; k0 < width; k0 += B) {
    uint32_t ent[B];
#pragma unroll
    for (int b = 0; b < B; ++b) ent[b] = (k0 + b < width) ? words[size_t(kSlice) * (k0 + b)] : 0xFFFFFFFFu;
    int grp[B];
#pragma unroll
    for (int b = 0; b < B; ++b) grp[b] = (live && k0 + b < width) ? ag.next() : -1;
    S xr[B][NV], xi[B][NV];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int idx = ent[b] != 0xFFFFFFFFu ? int(ent[b] & imask) : 0;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (CMODE == 0) {
          xr[b][v] = S(x[at(v, idx)]);
          xi[b][v] = 0;
        } else {
          xr[b][v] = S(x[2 * at(v, idx)]);
          xi[b][v] = S(x[2 * at(v, idx) + 1]);
        }
      }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      if (ent[b] == 0xFFFFFFFFu || grp[b] < 0) continue;
      const uint32_t e = ent[b];
      const int g = grp[b];
      const int off = int(s_meta[4 * g + 3]) + pattern_of(r, s_meta[4 * g + 1]);
      const int code = int((e >> a.img_bits) & 15u);
      const int ci = int((e >> (a.img_bits + 5)) & 31u);
      S scale = ((e >> (a.img_bits + 4)) & 1u) ? S(-1) : S(1);
      if (SYM) scale *= sq_r * S(s_crsq[code]);
      Cx<S> c;
      if (CMODE == 2) {
        c.re = S(s_amp[2 * off]) * scale;
        c.im = S(s_amp[2 * off + 1]) * scale;
      } else {
        c.re = S(s_amp[off]) * scale;
        c.im = 0;
      }
      if ((SYM == 1 || SYM == 3) && SHIFT) {
        const S pr = S(s_chr[2 * ci]), pi = S(s_chr[2 * ci + 1]);
        if (CMODE == 2) {
          const S re = c.re * pr - c.im * pi;
          c.im = c.re * pi + c.im * pr;
          c.re = re;
        } else {
          c.re *= pr;  // real characters (k = L/2): pi == 0
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (CMODE == 0) {
          acc[v].re += c.re * xr[b][v];
        } else {
          acc[v].re += c.re * xr[b][v] - c.im * xi[b][v];
          acc[v].im += c.re * xi[b][v] + c.im * xr[b][v];
        }
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (CMODE == 0) {
      y[at(v, row)] = acc[v].re;
    } else {
      y[2 * at(v, row)] = acc[v].re;
      y[2 * at(v, row) + 1] = acc[v].im;
    }
  }
}

template <typename S, int CMODE, int SYM, bool SHIFT, int NV, typename SX = S, bool IL = false>
__global__ void __launch_bounds__(kBlock) hv_stream_kernel(HvArgs a, const SX* __restrict__ x, S* __restrict__ y) {
  if (a.gate && *a.gate) return;
  stream_rows<S, CMODE, SYM, SHIFT, NV, SX, IL>(a, x, y, blockIdx.x);
}

// Scan points of one real family in lockstep: per point an amplitude table
// (gather form, real), a per-row diagonal, an output vector and a gate; up
// to 16 lanes of one precision (hv_scan_batch_kernel) or 16 FP32 and 8 FP64
// lanes together (hv_scan_mixed_kernel).
constexpr int kMaxScanBatch = 24;
struct ScanBatch {
  const double* amp[kMaxScanBatch];  // per lane: a table of n_amp entries
  const double* diag[kMaxScanBatch];
  double* y[kMaxScanBatch];
  const int* gate[kMaxScanBatch];
};

// y_b = H(c_b) x_b for the B lanes of an interleaved input x[row·B + b], on
// the compact stream (real amplitudes, no character phases). One lane per
// row as in stream_rows: the decode of an entry (target, amplitude entry,
// orbit code, sign) is shared by the B points, each point scales by its own
// table (LDS) and the B inputs of a target row are contiguous, so one
// gathered line serves all points.
template <typename SX, int B, int SYM>
__global__ void __launch_bounds__(kBlock) hv_scan_batch_kernel(HvArgs a, ScanBatch sb, const SX* __restrict__ x) {
  extern __shared__ double smem_d[];
  double* s_amp = smem_d;                // B tables of n_amp
  double* s_sq = s_amp + B * a.n_amp;    // 65
  double* s_crsq = s_sq + 65;            // 16
  bool on[B];
  bool any = false;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    on[b] = sb.y[b] != nullptr && !(sb.gate[b] && *sb.gate[b]);
    any = any || on[b];
  }
  if (!any) return;
  for (int i = threadIdx.x; i < B * a.n_amp; i += kBlock) {
    const int b = i / a.n_amp;
    s_amp[i] = on[b] ? sb.amp[b][i - b * a.n_amp] : 0.0;
  }
  for (int i = threadIdx.x; i < 65; i += kBlock) s_sq[i] = sqrt(double(i));
  for (int i = threadIdx.x; i < 16; i += kBlock) s_crsq[i] = a.code_rsq[i];
  __syncthreads();
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t slice = row / kSlice, lane = row % kSlice;
  const uint32_t n_slices = (a.n + kSlice - 1) / kSlice;
  if (slice >= n_slices) return;  // whole wavefronts leave together
  const bool live = row < a.n;
  const uint32_t base = a.slice_ptr[slice];
  const uint32_t width = (a.slice_ptr[slice + 1] - base) / kSlice;
  double sq_r = 1;
  if (SYM && live) sq_r = s_sq[a.orbits[row]];
  const uint32_t own = live ? row : 0u;
  double acc[B];
#pragma unroll
  for (int b = 0; b < B; ++b) acc[b] = on[b] ? sb.diag[b][own] * double(x[size_t(own) * B + b]) : 0.0;
  constexpr int E = 4;
  const uint32_t* words = a.stream + base + lane;
  const int so = a.c_b, sc = so + a.c_ob, ss = sc + a.c_cb;
  const uint32_t imask = (1u << a.c_b) - 1u;
  for (uint32_t k0 = 0; k0 < width; k0 += E) {
    uint32_t ent[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t w = words[size_t(kSlice) * min(k0 + e, width - 1)];
      ent[e] = k0 + e < width ? w : 0xFFFFFFFFu;
    }
    SX xv[E][B];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const size_t idx = ent[e] != 0xFFFFFFFFu ? size_t(ent[e] & imask) : 0;
#pragma unroll
      for (int b = 0; b < B; ++b) xv[e][b] = x[idx * B + b];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t w = ent[e];
      if (w == 0xFFFFFFFFu) continue;
      const int off = a.c_off_base + int(field(w, so, a.c_ob));
      double scale = field(w, ss, a.c_sb) ? -1.0 : 1.0;
      if (SYM) scale *= sq_r * s_crsq[field(w, sc, a.c_cb)];
#pragma unroll
      for (int b = 0; b < B; ++b) acc[b] += s_amp[b * a.n_amp + off] * scale * double(xv[e][b]);
    }
  }
  if (!live) return;
#pragma unroll
  for (int b = 0; b < B; ++b)
    if (on[b]) sb.y[b][row] = acc[b];
}

// Lanes per row of the on-the-fly scan action for B points.
constexpr int scan_walk_lanes(int b) { return b >= 2 ? 2 : 1; }

// The batched scan action of hv_scan_batch_kernel computed on the fly: per
// target the orbit walk and the state-directory lookup of the walk-mode
// hv_kernel, done once for the B points (T targets of a lane at a time, the
// walks sharing their rotation loops), then one gathered line of the
// interleaved inputs x[idx·B + b] serves all points. No index stream is
// read. A row's targets are taken in group order with the unreachable ones
// left out, which is the order of its stream words, and every factor is the
// stream kernel's (the amplitude entry, the reflection sign, sqrt(N_r) from
// the same LDS table and 1/sqrt(N_t) from the host table the orbit codes
// use), so both kernels give bit-identical results.
// LDS: B tables of the off-diagonal amplitudes (na entries each), sqrt(i),
// 1/sqrt(i) (i < 65), the off-diagonal group words and the directory cells.
template <typename SX, int B, int SYM>
__global__ void __launch_bounds__(kBlock) hv_scan_walk_kernel(HvArgs a, ScanBatch sb, const SX* __restrict__ x,
                                                              int off_base, const double* __restrict__ size_rsq) {
  extern __shared__ double smem_d[];
  const int na = a.n_amp - off_base;
  double* s_amp = smem_d;                                          // B * na
  double* s_sq = s_amp + B * na;                                   // 65
  double* s_rsq = s_sq + 65;                                       // 65
  uint32_t* s_meta = reinterpret_cast<uint32_t*>(s_rsq + 65);      // 4 * n_off
  uint32_t* s_cells = s_meta + 4 * a.n_off;                        // 2 * ncells (8-byte aligned)
  bool on[B];
  bool any = false;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    on[b] = sb.y[b] != nullptr && !(sb.gate[b] && *sb.gate[b]);
    any = any || on[b];
  }
  if (!any) return;
  for (int i = threadIdx.x; i < B * na; i += kBlock) {
    const int b = i / na;
    s_amp[i] = on[b] ? sb.amp[b][off_base + i - b * na] : 0.0;
  }
  for (int i = threadIdx.x; i < 65; i += kBlock) {
    s_sq[i] = sqrt(double(i));
    s_rsq[i] = size_rsq[i];
  }
  for (int i = threadIdx.x; i < 4 * a.n_off; i += kBlock) s_meta[i] = a.meta[4 * a.n_diag + i];
  for (int i = threadIdx.x; i < 2 * a.ncells; i += kBlock) s_cells[i] = a.dcells[i];
  __syncthreads();
  // P lanes per row (scan_walk_lanes): the lanes of a row take its targets
  // in turn for the walk and the lookup and share the results with lane
  // shuffles; each then gathers and accumulates its share of the points (BH
  // of them) for every target of the row, in the row's target order. Each
  // point's sum is thus the one-lane sum, term for term.
  constexpr int P = scan_walk_lanes(B);
  constexpr int BH = B / P;
  const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t row = gid / P;
  const int half = int(gid % P);
  if (row >= a.n) return;  // the lanes of a row leave together
  const int p0 = half * BH;
  const int lane0 = int(threadIdx.x & 63u) & ~(P - 1);  // the row's first lane in the wavefront
  bool onh[BH];
  const double* dg[BH];
  double* yp[BH];
#pragma unroll
  for (int q = 0; q < BH; ++q) {
    onh[q] = on[q];
    dg[q] = sb.diag[q];
    yp[q] = sb.y[q];
#pragma unroll
    for (int h = 1; h < P; ++h)
      if (half == h) {
        onh[q] = on[h * BH + q];
        dg[q] = sb.diag[h * BH + q];
        yp[q] = sb.y[h * BH + q];
      }
  }
  const uint32_t r = a.rstates[row];
  double sq_r = 1;
  if (SYM) sq_r = s_sq[a.orbits[row]];
  double acc[BH];
#pragma unroll
  for (int q = 0; q < BH; ++q) acc[q] = onh[q] ? dg[q][row] * double(x[size_t(row) * B + p0 + q]) : 0.0;
  const int nmax = a.max_sites;
  // Software-pipelined over batches (T targets walked per lane, P·T per
  // row; lane pairs with one target each measured fastest, against two
  // targets or four lanes, and a single lane with two): while the gathers
  // of one batch are in flight, the next batch is walked and its slots are
  // requested; then the pending batch is accumulated (in target order), the
  // new slots are resolved, swapped and their gathers issued. Per target the
  // swap carries the index and one word: amplitude entry | orbit size << 20
  // | reflection sign << 31.
  constexpr int T = P > 1 ? 1 : 2;
  constexpr int TT = P * T;
  SX xv[TT][BH];
  int pidx[TT], pmeta[TT];
  bool pending = false;
  auto accumulate = [&]() {
#pragma unroll
    for (int u = 0; u < TT; ++u) {
      if (pidx[u] < 0) continue;
      const int offs = pmeta[u] & 0xFFFFF;
      double scale = pmeta[u] < 0 ? -1.0 : 1.0;
      if (SYM) scale *= sq_r * s_rsq[(pmeta[u] >> 20) & 127];
#pragma unroll
      for (int q = 0; q < BH; ++q) acc[q] += s_amp[(p0 + q) * na + offs] * scale * double(xv[u][q]);
    }
  };
  for (int w0 = 0; w0 < a.n_off; w0 += 32) {
    uint32_t act = 0;
    const int gend = min(32, a.n_off - w0);
    for (int gi = 0; gi < gend; ++gi) {
      const int g = w0 + gi;
      const uint32_t packed = __builtin_amdgcn_readfirstlane(s_meta[4 * g + 1]);
      const uint32_t nz = __builtin_amdgcn_readfirstlane(s_meta[4 * g + 2]);
      act |= ((nz >> pattern_upto(r, packed, nmax)) & 1u) << gi;
    }
    while (act) {
      uint32_t tgt[T], key[T];
      bool in[T];
      int offs[T], cis[T], sgn[T], orb[T];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        // the pair's next P targets in order; this lane walks number `half`
        int sel = -1;
#pragma unroll
        for (int h = 0; h < P; ++h) {
          if (act) {
            const int gi = __ffs(act) - 1;
            act &= act - 1;
            if (h == half) sel = gi;
          }
        }
        in[t] = sel >= 0;
        tgt[t] = r;
        offs[t] = 0;
        if (sel >= 0) {
          const int g = w0 + sel;
          tgt[t] = r ^ s_meta[4 * g];
          offs[t] = int(s_meta[4 * g + 3]) + pattern_upto(r, s_meta[4 * g + 1], nmax) - off_base;
        }
      }
      if constexpr (SYM != 0) {
        walk_batch<SYM, false, T>(tgt, a, key, orb, cis, sgn);
      } else {
#pragma unroll
        for (int t = 0; t < T; ++t) {
          key[t] = tgt[t];
          orb[t] = 1;
          sgn[t] = 1;
        }
      }
      uint4 slot[T];
      dir_issue<T>(key, s_cells, a, slot);
      if (pending) accumulate();
      int idx[T];
      dir_resolve<T>(key, in, slot, idx);
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int m = offs[t] | (orb[t] << 20) | (sgn[t] < 0 ? int(0x80000000u) : 0);
#pragma unroll
        for (int h = 0; h < P; ++h) {
          pidx[P * t + h] = P > 1 ? __shfl(idx[t], lane0 + h, 64) : idx[t];
          pmeta[P * t + h] = P > 1 ? __shfl(m, lane0 + h, 64) : m;
        }
      }
      // the gathered halves of the lines (an unreachable target reads the
      // row's own line and adds nothing)
#pragma unroll
      for (int u = 0; u < TT; ++u) {
        const size_t at = pidx[u] >= 0 ? size_t(pidx[u]) : size_t(row);
#pragma unroll
        for (int q = 0; q < BH; ++q) xv[u][q] = x[at * B + p0 + q];
      }
      pending = true;
    }
  }
  if (pending) accumulate();
#pragma unroll
  for (int q = 0; q < BH; ++q)
    if (onh[q]) yp[q][row] = acc[q];
}

// The same action with lanes of both precisions in one launch: lanes
// b < W32 read the FP32 interleaved input x32[row·W32 + b], lanes
// W32 <= b < W32 + W64 the FP64 input x64[row·W64 + (b − W32)]. Windows of a
// continuous lockstep batch sit in different stages of the FP32 schedule (a
// window that has just taken its slot runs FP64 products first), and one
// pass over the index stream serves all of them. Every lane computes what
// hv_scan_batch_kernel computes for it (same terms, same order).
template <int W32, int W64, int SYM>
__global__ void __launch_bounds__(kBlock) hv_scan_mixed_kernel(HvArgs a, ScanBatch sb, const float* __restrict__ x32,
                                                               const double* __restrict__ x64) {
  constexpr int B = W32 + W64;
  extern __shared__ double smem_d[];
  double* s_amp = smem_d;                // B tables of n_amp
  double* s_sq = s_amp + B * a.n_amp;    // 65
  double* s_crsq = s_sq + 65;            // 16
  bool on[B];
  bool any = false;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    on[b] = sb.y[b] != nullptr && !(sb.gate[b] && *sb.gate[b]);
    any = any || on[b];
  }
  if (!any) return;
  for (int i = threadIdx.x; i < B * a.n_amp; i += kBlock) {
    const int b = i / a.n_amp;
    s_amp[i] = on[b] ? sb.amp[b][i - b * a.n_amp] : 0.0;
  }
  for (int i = threadIdx.x; i < 65; i += kBlock) s_sq[i] = sqrt(double(i));
  for (int i = threadIdx.x; i < 16; i += kBlock) s_crsq[i] = a.code_rsq[i];
  __syncthreads();
  const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t slice = row / kSlice, lane = row % kSlice;
  const uint32_t n_slices = (a.n + kSlice - 1) / kSlice;
  if (slice >= n_slices) return;  // whole wavefronts leave together
  const bool live = row < a.n;
  const uint32_t base = a.slice_ptr[slice];
  const uint32_t width = (a.slice_ptr[slice + 1] - base) / kSlice;
  double sq_r = 1;
  if (SYM && live) sq_r = s_sq[a.orbits[row]];
  const uint32_t own = live ? row : 0u;
  double acc[B];
#pragma unroll
  for (int b = 0; b < W32; ++b) acc[b] = on[b] ? sb.diag[b][own] * double(x32[size_t(own) * W32 + b]) : 0.0;
#pragma unroll
  for (int b = 0; b < W64; ++b)
    acc[W32 + b] = on[W32 + b] ? sb.diag[W32 + b][own] * x64[size_t(own) * W64 + b] : 0.0;
  constexpr int E = B > 12 ? 2 : 4;
  const uint32_t* words = a.stream + base + lane;
  const int so = a.c_b, sc = so + a.c_ob, ss = sc + a.c_cb;
  const uint32_t imask = (1u << a.c_b) - 1u;
  for (uint32_t k0 = 0; k0 < width; k0 += E) {
    uint32_t ent[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t w = words[size_t(kSlice) * min(k0 + e, width - 1)];
      ent[e] = k0 + e < width ? w : 0xFFFFFFFFu;
    }
    float xf[E][W32 > 0 ? W32 : 1];
    double xd[E][W64 > 0 ? W64 : 1];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const size_t idx = ent[e] != 0xFFFFFFFFu ? size_t(ent[e] & imask) : 0;
#pragma unroll
      for (int b = 0; b < W32; ++b) xf[e][b] = x32[idx * W32 + b];
#pragma unroll
      for (int b = 0; b < W64; ++b) xd[e][b] = x64[idx * W64 + b];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t w = ent[e];
      if (w == 0xFFFFFFFFu) continue;
      const int off = a.c_off_base + int(field(w, so, a.c_ob));
      double scale = field(w, ss, a.c_sb) ? -1.0 : 1.0;
      if (SYM) scale *= sq_r * s_crsq[field(w, sc, a.c_cb)];
#pragma unroll
      for (int b = 0; b < W32; ++b) acc[b] += s_amp[b * a.n_amp + off] * scale * double(xf[e][b]);
#pragma unroll
      for (int b = 0; b < W64; ++b) acc[W32 + b] += s_amp[(W32 + b) * a.n_amp + off] * scale * xd[e][b];
    }
  }
  if (!live) return;
#pragma unroll
  for (int b = 0; b < B; ++b)
    if (on[b]) sb.y[b][row] = acc[b];
}

// One sector of a sector-batched action: its tables and where its rows sit
// in the concatenated (complex) vectors.
struct BatchSector {
  HvArgs a;
  size_t offset;  // first row in the concatenated vectors
  int cmode;      // 1: real amplitudes, 2: complex amplitudes
  int shift;      // character phases needed (k != 0)
};

// y = H x on many sectors in one launch: block b applies the rows of block
// (item & 0xFFFF) of sector (item >> 16), item = blocks[b].
template <int SYM>
__global__ void __launch_bounds__(kBlock) hv_stream_batch_kernel(const BatchSector* __restrict__ sec,
                                                                 const uint32_t* __restrict__ blocks,
                                                                 const double* __restrict__ x, double* __restrict__ y) {
  const uint32_t item = blocks[blockIdx.x];
  const BatchSector& bs = sec[item >> 16];
  const uint32_t blk = item & 0xFFFFu;
  const double* xs = x + 2 * bs.offset;
  double* ys = y + 2 * bs.offset;
  if (bs.cmode == 2) {
    if (bs.shift) stream_rows<double, 2, SYM, true, 1>(bs.a, xs, ys, blk);
    else stream_rows<double, 2, SYM, false, 1>(bs.a, xs, ys, blk);
  } else {
    if (bs.shift) stream_rows<double, 1, SYM, true, 1>(bs.a, xs, ys, blk);
    else stream_rows<double, 1, SYM, false, 1>(bs.a, xs, ys, blk);
  }
}

// Per-row diagonal of the current amplitude tables (gather layout: one
// double per entry, or (re, im) when complex): out[row] = Σ re, out[n + row]
// = Σ im over the diagonal groups, in group order.
// pattern_of without branches: the five site fields of the packed word (a
// group has at most five sites; absent fields are zero) and a mask of the ns
// low bits
__device__ __forceinline__ int pattern_fixed(uint32_t r, uint32_t packed) {
  int pat = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) pat |= int(__builtin_amdgcn_ubfe(r, (packed >> (3 + 5 * i)) & 31u, 1)) << i;
  return pat & ((1 << (packed & 7u)) - 1);
}

template <typename ST>
__global__ void __launch_bounds__(kBlock) diag_kernel(HvArgs a, int complex_amp, double* out) {
  // The amplitude table and the diagonal groups' (sites, offset) words in
  // LDS. Every row reads n_diag entries at row-dependent offsets; the group
  // words arrive as vector registers, so the pattern extraction runs on the
  // vector ALU (the per-CU scalar unit would otherwise serialize it).
  __shared__ double s_amp[2 * kMaxAmps];
  __shared__ uint2 s_grp[kMaxGroups];
  const int namp = (complex_amp ? 2 : 1) * a.n_amp;
  for (int i = threadIdx.x; i < namp; i += kBlock) s_amp[i] = a.amp[i];
  if constexpr (sizeof(ST) == 4)
    for (int g = threadIdx.x; g < a.n_diag; g += kBlock) s_grp[g] = make_uint2(a.meta[4 * g + 1], a.meta[4 * g + 3]);
  __syncthreads();
  const int G = a.n_diag;
  if constexpr (sizeof(ST) == 8) {
    // wide states: the 64-bit group words, one group at a time
    for (size_t row = size_t(blockIdx.x) * kBlock + threadIdx.x; row < a.n; row += size_t(gridDim.x) * kBlock) {
      const uint64_t r = a.rstates64[row];
      double re = 0, im = 0;
      for (int g = 0; g < G; ++g) {
        const int o = int(a.meta64[4 * g + 3]) + pattern_of64(r, a.meta64[4 * g + 1]);
        if (complex_amp) {
          re += s_amp[2 * o];
          im += s_amp[2 * o + 1];
        } else {
          re += s_amp[o];
        }
      }
      out[row] = re;
      out[a.n + row] = im;
    }
    return;
  }
  for (size_t row = size_t(blockIdx.x) * kBlock + threadIdx.x; row < a.n; row += size_t(gridDim.x) * kBlock) {
    const uint32_t r = uint32_t(row_state<ST>(a, uint32_t(row)));
    double re0 = 0, re1 = 0, im0 = 0, im1 = 0;
    int g = 0;
    for (; g + 2 <= G; g += 2) {
      const uint2 m0 = s_grp[g], m1 = s_grp[g + 1];
      const int o0 = int(m0.y) + pattern_fixed(r, m0.x);
      const int o1 = int(m1.y) + pattern_fixed(r, m1.x);
      if (complex_amp) {
        re0 += s_amp[2 * o0];
        im0 += s_amp[2 * o0 + 1];
        re1 += s_amp[2 * o1];
        im1 += s_amp[2 * o1 + 1];
      } else {
        re0 += s_amp[o0];
        re1 += s_amp[o1];
      }
    }
    if (g < G) {
      const uint2 m0 = s_grp[g];
      const int o0 = int(m0.y) + pattern_fixed(r, m0.x);
      if (complex_amp) {
        re0 += s_amp[2 * o0];
        im0 += s_amp[2 * o0 + 1];
      } else {
        re0 += s_amp[o0];
      }
    }
    out[row] = re0 + re1;
    out[a.n + row] = im0 + im1;
  }
}

// Per-row diagonal of a family at coefficients c from the stored diagonals
// S_p of its parts (p = 0 the static part, then one per component; each S_p
// as re[n], im[n] in the gather form the H·v kernels read, the conjugate of
// the part's diagonal): the diagonal is linear in the coefficients, so the
// stored form at c is S_0 + Σ_p conj(c_p) S_p. With real part tables the S_p
// are real and only their real halves are read.
constexpr int kMaxDiagParts = 8;
struct DiagCoefs {
  int parts = 0;
  int real_parts = 0;   // every S_p real
  int complex_out = 0;  // the stored imaginary half is used (else zero)
  double re[kMaxDiagParts], im[kMaxDiagParts];  // conj(c_p), p >= 1
};

__global__ void __launch_bounds__(kBlock) diag_combine_kernel(const double* __restrict__ S, size_t n, DiagCoefs k,
                                                              double* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) {
    double re = S[i], im = k.real_parts ? 0.0 : S[n + i];
    for (int p = 1; p < k.parts; ++p) {
      const double sr = S[2 * n * size_t(p) + i];
      const double si = k.real_parts ? 0.0 : S[2 * n * size_t(p) + n + i];
      re += k.re[p] * sr - k.im[p] * si;
      im += k.re[p] * si + k.im[p] * sr;
    }
    out[i] = re;
    out[n + i] = k.complex_out ? im : 0.0;
  }
}

// Launch the 32-bit or 64-bit state instantiation of a build kernel (wide
// sectors hold 64-bit states).
#define QMBED_WIDTH_LAUNCH(basis, K, ...)                                \
  do {                                                                  \
    if ((basis).wide()) hipLaunchKernelGGL(K<uint64_t>, __VA_ARGS__);   \
    else hipLaunchKernelGGL(K<uint32_t>, __VA_ARGS__);                  \
  } while (0)

// The symmetry kind of the orbit walks (target_of): 0 none, 1 translation,
// 2 reflection, 3 both, 4 a general group
int sym_code(const SpinBasis1D& b) {
  const SectorGeometry& g = b.geometry();
  if (g.n_group) return 4;
  if (g.kmom >= 0 && g.parity) return 3;
  if (g.kmom >= 0) return 1;
  if (g.parity) return 2;
  return 0;
}

// The symmetry kind of the stream kernels, which read orbit codes, signs
// and character indices from the words: a general group has no separate
// sign, so its words decode like a translation sector's
int kernel_sym(const SpinBasis1D& b) { return b.general() ? 1 : sym_code(b); }

// the target words carry a character index
bool phased(const SpinBasis1D& b) {
  const SectorGeometry& g = b.geometry();
  return g.n_group ? g.n_chars > 1 : g.kmom > 0;
}

}  // namespace

// Device copies of the compiled flip groups and symmetry tables.
// amp: gather amplitudes conj(c·a) (real or (re, im)), read by the H·v
// kernels; famp: forward amplitudes c·a as (re, im), read by the column
// kernels.
struct Operator::DeviceTables {
  DeviceArray<uint32_t> meta;
  DeviceArray<uint64_t> meta64;  // wide sectors: [flip, sites, nonzero mask, offset] per group
  DeviceArray<double> amp;
  DeviceArray<double> famp;
  DeviceArray<double> chars;
  int n_diag = 0, n_off = 0, n_amp = 0;
  int max_sites = 0;
  bool complex_coef = false;  // amplitudes or characters complex
  bool hermitian = true;      // sector matrix Hermitian: gather form valid
  bool family = false;        // parameterized: amplitudes are re-evaluated
  int lpr = 1;                // lanes per row of the H·v kernels (autotuned)
  // image-table mode (autotuned against the orbit-walk mode)
  bool use_img = false;
  DeviceArray<uint32_t> img, hop, binom;
  DeviceArray<double> code_rsq;
  // 1/sqrt(i) for orbit sizes i < 65, computed on the host as the orbit
  // codes' factors are (the on-the-fly scan action), and whether the scan
  // action runs on the fly (QMBED_SCAN_HV=walk) rather than on the stream
  DeviceArray<double> size_rsq;
  int scan_walk = -1;
  int img_bits = 0;
  // compiled index stream (autotuned against both)
  bool use_stream = false;
  DeviceArray<uint32_t> stream, slice_ptr;
  // the stream is kept for the interleaved multi-column action even when
  // the single-vector autotuner prefers another mode
  bool use_stream_interleaved = false;
  // compact stream layout (field widths; see HvArgs), and the byte per word
  // of the split layout
  bool compact = false;
  DeviceArray<uint8_t> aux;
  int c_b = 0, c_ob = 0, c_cb = 0, c_sb = 0, c_off_base = 0;
  int n_codes = 1;
  // per-row diagonal of the current coefficients (re[n], im[n])
  DeviceArray<double> diag;
  // families: the stored per-row diagonal of every part (parts × (re[n],
  // im[n])), from which each evaluation combines the diagonal
  DeviceArray<double> part_diag;
  int diag_parts = 0;
  // stored sector matrix (a stored format was requested): CSR rows with
  // the reference's stored entries; built on demand for the current
  // evaluation (version counts the coefficient evaluations)
  bool use_csr = false;
  DeviceArray<uint64_t> csr_ptr;
  DeviceArray<uint32_t> csr_col;
  DeviceArray<double> csr_val;  // one double per entry (real) or (re, im)
  bool csr_complex = false;
  size_t csr_nnz = 0;
  long csr_version = -1;
  long version = 0;
  // pinned staging of the amplitude uploads and the event after their copies
  PinnedArray<double> stage;
  hipEvent_t staged = nullptr;
  ~DeviceTables() {
    if (staged) (void)hipEventDestroy(staged);
  }
  HvArgs args(const SpinBasis1D& b) const {
    HvArgs a{};
    const SectorGeometry& g = b.geometry();
    a.states = b.states();
    a.rstates = b.row_states();
    a.orbits = b.row_orbits();
    a.rev = b.reversed() ? 1 : 0;
    a.ngroup = g.n_group;
    a.gdest = b.group_dest();
    a.gflip = b.group_flip();
    a.gphase = b.group_phase();
    a.grot = b.group_rot();
    a.maxrep = g.max_rep ? 1 : 0;
    a.nchar = g.n_group ? g.n_chars : (g.kmom >= 0 ? g.L : 1);
    a.n = uint32_t(b.len());
    a.L = g.L;
    a.mask = uint32_t(g.mask);
    a.mask64 = g.mask;
    a.states64 = b.states64();
    a.rstates64 = b.row_states64();
    a.meta64 = meta64.data();
    a.parity = g.parity;
    a.n_div = g.n_div;
    for (int i = 0; i < 16; ++i) a.divisors[i] = g.divisors[i];
    a.dcells = b.directory_cells();
    a.dir = b.directory();
    a.dcb = b.directory_shift();
    a.ncells = b.wide() ? 0 : (1 << (g.L - a.dcb));
    a.lpr = lpr;
    uint32_t mask = 0;
    for (int i = 0; i < 32; i += lpr) mask |= 1u << i;
    a.lane_groups = mask;
    a.img = use_img ? img.data() : nullptr;
    a.img_bits = img_bits;
    a.up = (b.spec().up && b.spec().up_sectors.empty()) ? *b.spec().up : -1;
    a.hop = hop.data();
    a.binom = binom.data();
    a.code_rsq = code_rsq.data();
    a.meta = meta.data();
    a.amp = amp.data();
    a.n_diag = n_diag;
    a.n_off = n_off;
    a.max_sites = max_sites;
    a.n_amp = n_amp;
    a.chars = chars.data();
    a.nvec = 1;
    a.ld = b.len();
    a.gate = nullptr;
    a.stream = use_stream ? stream.data() : nullptr;
    a.slice_ptr = slice_ptr.data();
    a.compact = compact ? 1 : 0;
    a.aux = aux.size() ? aux.data() : nullptr;
    a.c_b = c_b;
    a.c_ob = c_ob;
    a.c_cb = c_cb;
    a.c_sb = c_sb;
    a.c_off_base = c_off_base;
    a.diag = diag.size() >= 2 * b.len() && b.len() > 0 ? diag.data() : nullptr;
    return a;
  }
};

// Hermiticity from the stored entries (two column passes, a radix sort and
// an entry-wise comparison); returns -1 when it cannot decide (an entry
// whose transposed partner is not stored, or too little device memory),
// else 1 (Hermitian) or 0.
static int stored_hermitian(const Operator& op, hipStream_t s) {
  const SpinBasis1D& b = op.basis();
  HvArgs a = op.tables_->args(b);
  const size_t n = a.n;
  DeviceArray<uint32_t> cnt(n);
  DeviceArray<int> flags(2);
  QMBED_HIP(hipMemsetAsync(flags.data(), 0, 2 * sizeof(int), s));
  const dim3 grid(unsigned((n + kBlock - 1) / kBlock));
  QMBED_WIDTH_LAUNCH(b, column_entries_kernel, grid, dim3(kBlock), 0, s, a, op.tables_->famp.data(), sym_code(b),
                     phased(b) ? 1 : 0, nullptr, nullptr, nullptr, nullptr, flags.data() + 1, cplx(0, 0), cnt.data());
  QMBED_HIP(hipGetLastError());
  std::vector<uint32_t> h(n);
  int of = 0;
  QMBED_HIP(hipMemcpyAsync(h.data(), cnt.data(), n * 4, hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipMemcpyAsync(&of, flags.data() + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  if (of) return -1;
  std::vector<uint64_t> ptr(n + 1, 0);
  for (size_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + h[i];
  const uint64_t nnz = ptr[n];
  if (nnz == 0) return 1;
  if (nnz >= (uint64_t(1) << 31)) return -1;
  size_t free_b = 0, total_b = 0;
  QMBED_HIP(hipMemGetInfo(&free_b, &total_b));
  if (nnz * 48 > free_b / 2) return -1;
  DeviceArray<uint64_t> cptr(n + 1);
  QMBED_HIP(hipMemcpyAsync(cptr.data(), ptr.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
  DeviceArray<uint32_t> rows(nnz), ccols(nnz), keys(nnz), perm_in(nnz), perm(nnz);
  DeviceArray<double2> vals(nnz);
  QMBED_WIDTH_LAUNCH(b, column_fill_kernel, grid, dim3(kBlock), 0, s, a, op.tables_->famp.data(), sym_code(b),
                     phased(b) ? 1 : 0, cptr.data(), rows.data(), vals.data(), flags.data() + 1);
  hipLaunchKernelGGL(csc_columns_kernel, grid, dim3(kBlock), 0, s, cptr.data(), uint32_t(n), ccols.data());
  std::vector<uint32_t> iota(nnz);
  std::iota(iota.begin(), iota.end(), 0u);
  QMBED_HIP(hipMemcpyAsync(perm_in.data(), iota.data(), nnz * 4, hipMemcpyHostToDevice, s));
  int bits = 1;
  while ((size_t(1) << bits) < n) ++bits;
  size_t tmp_bytes = 0;
  QMBED_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, rows.data(), keys.data(), perm_in.data(),
                                               perm.data(), int(nnz), 0, bits, s));
  DeviceArray<uint8_t> tmp(std::max<size_t>(tmp_bytes, 1));
  QMBED_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.data(), tmp_bytes, rows.data(), keys.data(), perm_in.data(),
                                               perm.data(), int(nnz), 0, bits, s));
  QMBED_HIP(hipMemsetAsync(flags.data(), 0, 2 * sizeof(int), s));
  hipLaunchKernelGGL(herm_compare_kernel, dim3(unsigned((nnz + 255) / 256)), dim3(256), 0, s, keys.data(),
                     perm.data(), ccols.data(), rows.data(), vals.data(), nnz, flags.data());
  QMBED_HIP(hipGetLastError());
  int hf[2] = {0, 0};
  QMBED_HIP(hipMemcpyAsync(hf, flags.data(), 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  if (hf[1]) return -1;
  return hf[0] ? 0 : 1;
}

static bool sector_matrix_hermitian(const Operator& op, hipStream_t s) {
  const SpinBasis1D& b = op.basis();
  HvArgs a = op.tables_->args(b);
  if (a.n == 0) return true;
  // the entry-wise check from the stored entries when it can decide; the
  // per-entry re-evaluation below otherwise (the reference's rule either way)
  const int fast = stored_hermitian(op, s);
  if (fast >= 0) return fast == 1;
  DeviceArray<int> flags(2);
  QMBED_HIP(hipMemsetAsync(flags.data(), 0, 2 * sizeof(int), s));
  QMBED_WIDTH_LAUNCH(b, hermiticity_kernel, dim3((a.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a,
                     op.tables_->famp.data(), sym_code(b), phased(b) ? 1 : 0, flags.data());
  QMBED_HIP(hipGetLastError());
  int h[2] = {0, 0};
  QMBED_HIP(hipMemcpyAsync(h, flags.data(), 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  if (h[1]) throw unsupported_backend("a column has more distinct entries than the Hermiticity check holds");
  return h[0] == 0;
}

static void tune_lanes(Operator& op);

// Amplitude tables of the current groups: the gather-form table conj(c·a)
// (one double per entry for real coefficients, (re, im) otherwise) and the
// forward table c·a as (re, im). For families the device arrays are sized
// for (re, im) so that an evaluation may switch between the layouts.
static void upload_amplitudes(Operator& op, const std::vector<FlipGroup>& diag, const std::vector<FlipGroup>& off,
                              hipStream_t s, const std::vector<cplx>* coefs = nullptr, bool real_parts = false) {
  auto& t = *op.tables_;
  std::vector<double> amp, famp;
  for (auto* v : {&diag, &off}) {
    for (auto& g : *v) {
      for (auto& c : g.amp) {
        amp.push_back(c.real());
        if (t.complex_coef) amp.push_back(-c.imag());  // gather form reads conj(c·a)
        famp.push_back(c.real());
        famp.push_back(c.imag());
      }
    }
  }
  const size_t cap = t.family ? 2 * size_t(std::max(t.n_amp, 1)) : std::max<size_t>(amp.size(), 2);
  if (t.amp.size() < cap) t.amp.resize(cap);
  if (t.famp.size() < std::max<size_t>(famp.size(), 2)) t.famp.resize(std::max<size_t>(famp.size(), 2));
  // staged through pinned memory, so the copies are asynchronous; the
  // staging is rewritten only after the previous upload's copies are done
  if (t.staged) QMBED_HIP(hipEventSynchronize(t.staged));
  t.stage.ensure(amp.size() + famp.size());
  std::copy(amp.begin(), amp.end(), t.stage.p);
  std::copy(famp.begin(), famp.end(), t.stage.p + amp.size());
  if (!amp.empty()) QMBED_HIP(hipMemcpyAsync(t.amp.data(), t.stage.p, amp.size() * 8, hipMemcpyHostToDevice, s));
  if (!famp.empty())
    QMBED_HIP(hipMemcpyAsync(t.famp.data(), t.stage.p + amp.size(), famp.size() * 8, hipMemcpyHostToDevice, s));
  // the per-row diagonal read by the stream kernel (one pass per evaluation
  // instead of a walk over the diagonal groups in every H·v)
  const SpinBasis1D& b = op.basis();
  const size_t n = b.len();
  if (n > 0 && n <= (size_t(1) << 28) && coefs && t.diag_parts == int(coefs->size()) + 1) {
    if (t.diag.size() != 2 * n) t.diag.resize(2 * n);
    DiagCoefs k;
    k.parts = t.diag_parts;
    k.real_parts = real_parts ? 1 : 0;
    k.complex_out = t.complex_coef ? 1 : 0;
    k.re[0] = 1.0;
    k.im[0] = 0.0;
    for (size_t p = 0; p < coefs->size(); ++p) {
      k.re[p + 1] = (*coefs)[p].real();
      k.im[p + 1] = -(*coefs)[p].imag();
    }
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(diag_combine_kernel, dim3(unsigned(blocks)), dim3(kBlock), 0, s, t.part_diag.data(), n, k,
                       t.diag.data());
    QMBED_HIP(hipGetLastError());
  } else if (n > 0 && n <= (size_t(1) << 28)) {
    if (t.diag.size() != 2 * n) t.diag.resize(2 * n);
    HvArgs a = t.args(b);
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
    QMBED_WIDTH_LAUNCH(b, diag_kernel, dim3(unsigned(blocks)), dim3(kBlock), 0, s, a, t.complex_coef ? 1 : 0,
                       t.diag.data());
    QMBED_HIP(hipGetLastError());
  }
  if (!t.staged) QMBED_HIP(hipEventCreateWithFlags(&t.staged, hipEventDisableTiming));
  QMBED_HIP(hipEventRecord(t.staged, s));
}

// The sector matrix of the current tables is Hermitian (checked on the
// device) — with a family, of each part in turn when `part` >= 0.
static bool current_hermitian(Operator& op, hipStream_t s) { return sector_matrix_hermitian(op, s); }

std::shared_ptr<Operator> OperatorBuilder::build(MatrixFormat format) const {
  if (!basis_) throw invalid_options("operator builder has no basis");
  auto op = std::shared_ptr<Operator>(new Operator());
  op->basis_ = basis_;
  op->format_ = format;
  const bool family = !components_.empty();
  if (family) {
    // QuantumOperator::new (src/operator/mod.rs:2717-2745)
    std::set<std::string> names;
    for (const auto& c : components_) {
      if (c.name.empty() || !names.insert(c.name).second)
        throw invalid_options("component names must be nonempty and unique");
      if (c.has_default && (!std::isfinite(c.default_value.real()) || !std::isfinite(c.default_value.imag())))
        throw invalid_options("component defaults must be finite");
    }
  }
  bool real_amp = true;
  if (family) {
    std::vector<std::vector<OperatorSpec>> parts{terms_};
    for (const auto& c : components_) parts.push_back(c.terms);
    compile_family_groups(basis_->spec(), parts, checks_, basis_->len() > 0, op->diag_, op->off_, real_amp);
  } else {
    compile_groups(basis_->spec(), terms_, checks_, basis_->len() > 0, op->diag_, op->off_, real_amp);
  }
  const SectorGeometry& geo = basis_->geometry();
  op->real_ = real_amp && basis_->real_characters();
  op->parts_real_ = real_amp;
  auto t = std::make_shared<Operator::DeviceTables>();
  t->family = family;
  t->n_diag = int(op->diag_.size());
  t->n_off = int(op->off_.size());
  if (t->n_diag + t->n_off > kMaxGroups)
    throw unsupported_backend("the operator has more distinct local actions than the H·v kernel tables hold");
  t->complex_coef = !op->real_;
  std::vector<uint32_t> meta;
  std::vector<uint64_t> meta64;
  int n_amp = 0;
  for (auto* v : {&op->diag_, &op->off_}) {
    for (auto& g : *v) {
      uint32_t packed = uint32_t(g.nsites);
      uint64_t packed64 = uint64_t(g.nsites);
      for (int i = 0; i < g.nsites; ++i) {
        packed |= uint32_t(g.sites[i] & 31) << (3 + 5 * i);
        packed64 |= uint64_t(g.sites[i]) << (3 + 6 * i);
      }
      t->max_sites = std::max(t->max_sites, int(g.nsites));
      // nonzero patterns: of the table, or of any part of a family (the
      // structure must not depend on the coefficients)
      uint32_t nz = 0;
      for (size_t p = 0; p < g.amp.size(); ++p) {
        bool on = g.amp[p] != cplx(0, 0);
        for (auto& part : g.parts) on = on || part[p] != cplx(0, 0);
        if (on) nz |= 1u << p;
      }
      meta.push_back(uint32_t(g.flip));
      meta.push_back(packed);
      meta.push_back(nz);
      meta.push_back(uint32_t(n_amp));
      meta64.push_back(g.flip);
      meta64.push_back(packed64);
      meta64.push_back(nz);
      meta64.push_back(uint64_t(n_amp));
      n_amp += int(g.amp.size());
    }
  }
  if (n_amp > kMaxAmps) throw unsupported_backend("the operator's amplitude tables exceed the H·v kernel's LDS budget");
  t->n_amp = n_amp;
  std::vector<double> chars(2 * kMaxChars, 0.0);
  if (geo.n_group) {
    const auto& cv = basis_->characters();
    for (size_t i = 0; i < cv.size() && i < size_t(kMaxChars); ++i) {
      chars[2 * i] = cv[i].real();
      chars[2 * i + 1] = cv[i].imag();
    }
  } else if (geo.kmom >= 0) {
    const double tau = 6.283185307179586;
    for (int s = 0; s < geo.L; ++s) {
      double angle = -tau * double(geo.kmom * s) / double(geo.L);
      if (2 * geo.kmom == geo.L) {
        chars[2 * s] = (s & 1) ? -1.0 : 1.0;  // exact ±1 at k = L/2
        chars[2 * s + 1] = 0.0;
      } else if (geo.kmom == 0) {
        chars[2 * s] = 1.0;
      } else {
        chars[2 * s] = std::cos(angle);
        chars[2 * s + 1] = std::sin(angle);
      }
    }
  } else {
    chars[0] = 1.0;
  }
  t->meta.resize(std::max<size_t>(meta.size(), 4));
  t->chars.resize(chars.size());
  if (!meta.empty()) QMBED_HIP(hipMemcpy(t->meta.data(), meta.data(), meta.size() * 4, hipMemcpyHostToDevice));
  if (basis_->wide()) {
    t->meta64.resize(std::max<size_t>(meta64.size(), 4));
    if (!meta64.empty())
      QMBED_HIP(hipMemcpy(t->meta64.data(), meta64.data(), meta64.size() * 8, hipMemcpyHostToDevice));
  }
  QMBED_HIP(hipMemcpy(t->chars.data(), chars.data(), chars.size() * 8, hipMemcpyHostToDevice));
  op->tables_ = t;
  if (!family) {
    upload_amplitudes(*op, op->diag_, op->off_, nullptr);
    // The gather-form kernels read row r as conj(column r), which is the
    // sector matrix exactly when it is Hermitian; the reference's assembly
    // checks exactly that (1e-12 per entry) unless checks are disabled.
    if (basis_->len() > 0) {
      if (checks_.hermiticity) {
        if (!current_hermitian(*op, nullptr)) throw non_hermitian();
      } else {
        t->hermitian = current_hermitian(*op, nullptr);
        if (!t->hermitian) op->real_ = false;
      }
      if (t->hermitian) {
        tune_lanes(*op);
        op->tuned_ = true;
      }
    }
    return op;
  }
  // Family: the static part is checked like a plain operator (the reference
  // materializes it with the model's checks, src/interop.rs:1277-1280); each
  // component's Hermiticity is recorded (components are built unchecked,
  // :1256-1260) so that real-coefficient sums of Hermitian parts need no
  // check when evaluated.
  op->components_ = components_;
  const size_t nparts = components_.size() + 1;
  op->part_hermitian_.assign(nparts, true);
  if (basis_->len() > 0) {
    // the parts' stored diagonals, kept when they fit (at most 8 parts and
    // 8 GiB), so that an evaluation combines them instead of walking the
    // diagonal groups of every row again
    const size_t n = basis_->len();
    const bool keep = nparts <= size_t(kMaxDiagParts) && n <= (size_t(1) << 28) &&
                      nparts * 2 * n * sizeof(double) <= (size_t(8) << 30);
    if (keep) t->part_diag.resize(nparts * 2 * n);
    for (size_t part = 0; part < nparts; ++part) {
      std::vector<FlipGroup> d = op->diag_, o = op->off_;
      for (auto* v : {&d, &o})
        for (auto& g : *v) g.amp = g.parts[part];
      t->complex_coef = !(op->parts_real_ && basis_->real_characters());
      upload_amplitudes(*op, d, o, nullptr);
      if (keep)
        QMBED_HIP(hipMemcpy(t->part_diag.data() + part * 2 * n, t->diag.data(), 2 * n * sizeof(double),
                            hipMemcpyDeviceToDevice));
      op->part_hermitian_[part] = current_hermitian(*op, nullptr);
    }
    if (keep) t->diag_parts = int(nparts);
    if (checks_.hermiticity && !op->part_hermitian_[0]) throw non_hermitian();
  }
  // initial evaluation: defaults, and zero for required components (an
  // evaluation with their values must come before any use)
  std::vector<cplx> c(components_.size(), cplx(0, 0));
  for (size_t p = 0; p < components_.size(); ++p)
    if (components_[p].has_default) c[p] = components_[p].default_value;
  op->evaluate_coefficients(c, nullptr);
  return op;
}

std::vector<cplx> Operator::resolve_coefficients(const std::vector<std::pair<std::string, cplx>>& parameters) const {
  // src/operator/mod.rs:2775-2812 (and src/interop.rs:964-973 without components)
  for (const auto& kv : parameters) {
    bool known = false;
    for (const auto& c : components_) known = known || c.name == kv.first;
    if (!known) throw invalid_options("unknown operator parameter " + debug_string(kv.first));
  }
  std::vector<cplx> out;
  for (const auto& c : components_) {
    const cplx* found = nullptr;
    for (const auto& kv : parameters)
      if (kv.first == c.name) found = &kv.second;
    if (!found && !c.has_default)
      throw invalid_options("missing required operator parameter " + debug_string(c.name));
    const cplx v = found ? *found : c.default_value;
    if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
      throw invalid_options("operator parameter " + debug_string(c.name) + " must be finite");
    out.push_back(v);
  }
  return out;
}

void Operator::evaluate_coefficients(const std::vector<cplx>& c, hipStream_t s) {
  if (!parameterized()) {
    if (!c.empty()) throw dimension_mismatch("received " + std::to_string(c.size()) + " operator coefficients for 0 components");
    return;
  }
  if (c.size() != components_.size())
    throw dimension_mismatch("received " + std::to_string(c.size()) + " operator coefficients for " +
                             std::to_string(components_.size()) + " components");
  for (const auto& v : c)
    if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
      throw invalid_options("operator coefficients must be finite");
  bool real_c = true;
  for (const auto& v : c) real_c = real_c && v.imag() == 0.0;
  for (auto* v : {&diag_, &off_}) {
    for (auto& g : *v) {
      for (size_t pat = 0; pat < g.amp.size(); ++pat) {
        cplx a = g.parts[0][pat];
        for (size_t p = 0; p < c.size(); ++p) a += c[p] * g.parts[p + 1][pat];
        g.amp[pat] = a;
      }
    }
  }
  coefficients_ = c;
  auto& t = *tables_;
  ++t.version;
  real_ = parts_real_ && real_c && basis_->real_characters();
  t.complex_coef = !real_;
  upload_amplitudes(*this, diag_, off_, s, &c, parts_real_);
  if (basis_->len() == 0) return;
  bool all_parts = real_c;
  for (bool h : part_hermitian_) all_parts = all_parts && h;
  t.hermitian = all_parts ? true : current_hermitian(*this, s);
  if (!t.hermitian) real_ = false;
  if (t.hermitian && !tuned_) {
    tune_lanes(*this);
    tuned_ = true;
  }
}

bool Operator::hermitian() const { return tables_->hermitian; }

namespace {

size_t hv_smem(const HvArgs& a, int cmode, bool img) {
  const size_t G = size_t(a.n_diag + a.n_off);
  return sizeof(double) * ((cmode == 2 ? 2 : 1) * size_t(a.n_amp) + 2 * kMaxChars + 130 + 16) +
         sizeof(uint32_t) * (4 * G + G + (img ? 1024 : 2 * size_t(a.ncells)));
}

template <typename S, int CMODE, int NV, bool IMG>
void launch_hv_mode(const HvArgs& a, int sym, bool shift, const S* x, S* y, hipStream_t s) {
  const size_t lanes = size_t(a.n) * size_t(a.lpr);
  dim3 grid(unsigned((lanes + kBlock - 1) / kBlock)), block(kBlock);
  size_t sm = hv_smem(a, CMODE, IMG);
#define QMBED_HV_CASE(SYMV, SH) \
  hipLaunchKernelGGL((hv_kernel<S, CMODE, SYMV, SH, NV, IMG>), grid, block, sm, s, a, x, y);
  switch (sym) {
    case 0: QMBED_HV_CASE(0, false); break;
    case 1:
      if (shift) {
        QMBED_HV_CASE(1, true);
      } else {
        QMBED_HV_CASE(1, false);
      }
      break;
    case 2: QMBED_HV_CASE(2, false); break;
    default:
      if (shift) {
        QMBED_HV_CASE(3, true);
      } else {
        QMBED_HV_CASE(3, false);
      }
      break;
  }
#undef QMBED_HV_CASE
  QMBED_HIP(hipGetLastError());
}

template <typename S, int CMODE, int NV, typename SX = S, bool IL = false>
void launch_hv_stream(const HvArgs& a, int sym, bool shift, const SX* x, S* y, hipStream_t s) {
  const size_t n_slices = (size_t(a.n) + kSlice - 1) / kSlice;
  dim3 grid(unsigned((n_slices * kSlice + kBlock - 1) / kBlock)), block(kBlock);
  const size_t G = size_t(a.n_diag + a.n_off);
  const size_t sm =
      sizeof(double) * ((CMODE == 2 ? 2 : 1) * size_t(a.n_amp) + 2 * kMaxChars + 65 + 16) + sizeof(uint32_t) * 4 * G;
#define QMBED_HVS_CASE(SYMV, SH) \
  hipLaunchKernelGGL((hv_stream_kernel<S, CMODE, SYMV, SH, NV, SX, IL>), grid, block, sm, s, a, x, y);
  switch (sym) {
    case 0: QMBED_HVS_CASE(0, false); break;
    case 1:
      if (shift) {
        QMBED_HVS_CASE(1, true);
      } else {
        QMBED_HVS_CASE(1, false);
      }
      break;
    case 2: QMBED_HVS_CASE(2, false); break;
    default:
      if (shift) {
        QMBED_HVS_CASE(3, true);
      } else {
        QMBED_HVS_CASE(3, false);
      }
      break;
  }
#undef QMBED_HVS_CASE
  QMBED_HIP(hipGetLastError());
}

template <typename S, int CMODE, int NV>
void launch_hv(const HvArgs& a, int sym, bool shift, const S* x, S* y, hipStream_t s) {
  if (a.n == 0) return;
  if (!a.stream && (a.ngroup || a.rev || a.states64))
    throw internal_state("general, reversed and wide sectors are applied through the compiled stream only");
  if (a.stream) launch_hv_stream<S, CMODE, NV>(a, sym, shift, x, y, s);
  else if (a.img) launch_hv_mode<S, CMODE, NV, true>(a, sym, shift, x, y, s);
  else launch_hv_mode<S, CMODE, NV, false>(a, sym, shift, x, y, s);
}

}  // namespace

// Autotuner: lanes per row. Small sectors leave most of the 256 CUs idle
// with one lane per row; splitting a row's flip groups over 2-8 lanes (and
// reducing with cross-lane shuffles) restores parallelism. The FP64 kernel is
// timed with HIP events for each candidate and the fastest is kept.
// Orbit codes of the packed target words (image table and stream): every
// possible orbit size gets a 4-bit code, 1/sqrt(size) per code goes to the
// device, and the representative index takes the low img_bits bits. Returns
// false when the sector does not fit the packing.
static bool prepare_codes(Operator& op, std::vector<uint8_t>& code) {
  const SpinBasis1D& b = op.basis();
  const SectorGeometry& g = b.geometry();
  auto& t = *op.tables_;
  std::vector<int> sizes;
  for (int d = 1; d <= g.L; ++d)
    if (g.L % d == 0) {
      sizes.push_back(d);
      if (g.parity) sizes.push_back(2 * d);
    }
  if (!b.symmetric()) sizes = {1};
  if (g.kmom < 0 && g.parity) sizes = {1, 2};
  if (b.general()) {  // the orbit sizes the sector holds
    sizes.clear();
    std::vector<uint8_t> o = b.orbits_host();
    std::vector<char> seen(256, 0);
    for (uint8_t v : o) seen[v] = 1;
    for (int v = 1; v < 256; ++v)
      if (seen[size_t(v)]) sizes.push_back(v);
    if (sizes.empty()) sizes = {1};
  }
  std::sort(sizes.begin(), sizes.end());
  sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
  if (sizes.size() > 16) return false;
  code.assign(256, 0);
  std::vector<double> crsq(16, 0.0);
  for (size_t i = 0; i < sizes.size(); ++i) {
    code[size_t(sizes[i])] = uint8_t(i);
    crsq[i] = 1.0 / std::sqrt(double(sizes[i]));
  }
  int bits = 1;
  while ((size_t(1) << bits) < b.len()) ++bits;
  t.code_rsq.resize(16);
  QMBED_HIP(hipMemcpy(t.code_rsq.data(), crsq.data(), 16 * 8, hipMemcpyHostToDevice));
  t.img_bits = bits;
  t.n_codes = int(sizes.size());
  return true;
}

// Build the compiled index stream (count pass, slice offsets on the host,
// fill pass). Returns false when it does not fit the memory budget.
static int bits_for(size_t values) {
  int bits = 0;
  while ((size_t(1) << bits) < values) ++bits;
  return bits;
}

static bool build_stream(Operator& op, hipStream_t s) {
  const SpinBasis1D& b = op.basis();
  auto& t = *op.tables_;
  bool split = false;
  std::vector<uint8_t> code;
  if (!prepare_codes(op, code)) return false;
  const size_t n = b.len();
  // compact words when every field fits in 32 bits: index (all ones is never
  // an index), amplitude entry among the off-diagonal tables, orbit code,
  // reflection sign (parity sectors), χ-index (k != 0)
  {
    const SectorGeometry& g = b.geometry();
    int base = 0;  // the off-diagonal tables follow the diagonal ones
    for (auto& grp : op.diagonal_groups()) base += int(grp.amp.size());
    t.c_off_base = base;
    t.c_b = bits_for(n + 1);
    t.c_ob = bits_for(size_t(std::max(t.n_amp - base, 1)));
    t.c_cb = bits_for(size_t(t.n_codes));
    t.c_sb = g.parity ? 1 : 0;
    const int chb = phased(b) ? bits_for(size_t(g.n_group ? g.n_chars : g.L)) : 0;
    // QMBED_HV_STREAM_FORMAT=groups keeps the group-walk words everywhere
    const char* fmt = std::getenv("QMBED_HV_STREAM_FORMAT");
    const bool force_groups = fmt && std::string(fmt) == "groups";
    t.compact = !force_groups && t.c_b + t.c_ob + t.c_cb + t.c_sb + chb <= 32;
    // the split layout when the fields above the entry fit one byte
    split = !force_groups && !t.compact && t.c_b + t.c_ob <= 32 && t.c_cb + t.c_sb + chb <= 8;
    if (split) t.compact = true;
    // a general group's words are compact only (its character indices may
    // exceed the group-walk word's 5 bits), and so are a wide sector's (its
    // products read no states) when its per-row diagonal exists
    if (b.general() && !t.compact) return false;
    if (b.wide() && !(t.compact && t.diag.size() >= 2 * n)) return false;
    // the group-walk word: index, 4 code bits, sign, 5 χ-index bits
    if (!t.compact && t.img_bits + 10 > 32) return false;
  }
  const uint32_t n_slices = uint32_t((n + kSlice - 1) / kSlice);
  HvArgs a = t.args(b);
  DeviceArray<uint32_t> width(n_slices);
  const unsigned grid = unsigned((size_t(n_slices) * kSlice + kBlock - 1) / kBlock);
  QMBED_WIDTH_LAUNCH(b, stream_count_kernel, dim3(grid), dim3(kBlock), 0, s, a, n_slices, width.data());
  QMBED_HIP(hipGetLastError());
  std::vector<uint32_t> w(n_slices);
  QMBED_HIP(hipMemcpyAsync(w.data(), width.data(), size_t(n_slices) * 4, hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> ptr(size_t(n_slices) + 1, 0);
  uint64_t total = 0;
  for (uint32_t sl = 0; sl < n_slices; ++sl) {
    ptr[sl] = uint32_t(total);
    total += uint64_t(kSlice) * w[sl];
    if (total >= (uint64_t(1) << 32)) return false;
  }
  ptr[n_slices] = uint32_t(total);
  size_t free_b = 0, total_b = 0;
  QMBED_HIP(hipMemGetInfo(&free_b, &total_b));
  if (total * (split ? 5 : 4) > std::min<uint64_t>(uint64_t(16) << 30, free_b / 4)) return false;
  t.slice_ptr.resize(ptr.size());
  QMBED_HIP(hipMemcpy(t.slice_ptr.data(), ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice));
  t.stream.resize(std::max<uint64_t>(total, 1));
  if (split) t.aux.resize(std::max<uint64_t>(total, 1));
  else t.aux.release();
  DeviceArray<uint8_t> dcode(256);
  QMBED_HIP(hipMemcpy(dcode.data(), code.data(), 256, hipMemcpyHostToDevice));
  a = t.args(b);
  QMBED_WIDTH_LAUNCH(b, stream_fill_kernel, dim3(grid), dim3(kBlock), 0, s, a, sym_code(b), phased(b) ? 1 : 0, n_slices,
                     dcode.data(), t.slice_ptr.data(), t.stream.data(), split ? t.aux.data() : nullptr);
  QMBED_HIP(hipGetLastError());
  QMBED_HIP(hipStreamSynchronize(s));
  return true;
}

// Build the image table of the operator's sector (one pass of the orbit walk
// and sorted-table lookup over every parent state). Returns false when the
// sector does not fit the packed format or the memory budget.
static bool build_image_table(Operator& op, hipStream_t s) {
  const SpinBasis1D& b = op.basis();
  const SpinBasisSpec& spec = b.spec();
  const SectorGeometry& g = b.geometry();
  auto& t = *op.tables_;
  // host tables: binomials, hops, orbit codes
  std::vector<uint32_t> binom(1024, 0);
  for (int n = 0; n < 32; ++n) {
    binom[n * 32] = 1;
    for (int k = 1; k < 32; ++k) binom[n * 32 + k] = n == 0 ? 0 : binom[(n - 1) * 32 + k - 1] + binom[(n - 1) * 32 + k];
  }
  std::vector<uint32_t> hop;
  for (auto* v : {&op.diagonal_groups(), &op.offdiagonal_groups()})
    for (auto& grp : *v) {
      uint32_t h = 0;
      if (grp.flip && __builtin_popcountll(grp.flip) == 2) {
        const int p = __builtin_ctzll(grp.flip), q = 63 - __builtin_clzll(grp.flip);
        // a two-site hop: amplitudes only where exactly one of p, q is up
        bool hopping = true;
        for (size_t pat = 0; pat < grp.amp.size(); ++pat) {
          if (grp.amp[pat] == cplx(0, 0)) continue;
          int bp = -1, bq = -1;
          for (int i = 0; i < grp.nsites; ++i) {
            if (grp.sites[i] == p) bp = int((pat >> i) & 1);
            if (grp.sites[i] == q) bq = int((pat >> i) & 1);
          }
          if (bp < 0 || bq < 0 || bp == bq) hopping = false;
        }
        if (hopping) h = uint32_t(p) | (uint32_t(q) << 8) | (1u << 16);
      }
      hop.push_back(h);
    }
  // the group walk of a general sector and the row reversal are served by
  // the compiled stream only
  if (b.general() || b.reversed() || b.wide()) return false;
  std::vector<uint8_t> code;
  if (!prepare_codes(op, code)) return false;
  if (t.img_bits + 10 > 32) return false;  // index, code, sign and χ-index in one word
  const bool fixed_up = spec.up.has_value() && spec.up_sectors.empty();
  uint64_t parents = 0, popmask = ~0ull;
  if (fixed_up) {
    parents = 1;
    for (int i = 1; i <= *spec.up; ++i) parents = parents * uint64_t(g.L - *spec.up + i) / uint64_t(i);
  } else {
    parents = uint64_t(1) << g.L;
    if (!spec.up_sectors.empty()) {
      popmask = 0;
      for (int u : spec.up_sectors) popmask |= 1ull << u;
    }
  }
  if (parents >= (uint64_t(1) << 32)) return false;
  size_t free_b = 0, total_b = 0;
  QMBED_HIP(hipMemGetInfo(&free_b, &total_b));
  if (parents * 4 > std::min<uint64_t>(uint64_t(16) << 30, free_b / 4)) return false;
  t.binom.resize(1024);
  t.hop.resize(std::max<size_t>(hop.size(), 1));
  DeviceArray<uint8_t> dcode(256);
  QMBED_HIP(hipMemcpy(t.binom.data(), binom.data(), 4096, hipMemcpyHostToDevice));
  if (!hop.empty()) QMBED_HIP(hipMemcpy(t.hop.data(), hop.data(), hop.size() * 4, hipMemcpyHostToDevice));
  QMBED_HIP(hipMemcpy(dcode.data(), code.data(), 256, hipMemcpyHostToDevice));
  t.img.resize(parents);
  HvArgs a = t.args(b);
  const uint64_t threads = (parents + kImgChunk - 1) / kImgChunk;
  hipLaunchKernelGGL(image_kernel, dim3(unsigned((threads + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, a, sym_code(b),
                     phased(b) ? 1 : 0, popmask, uint32_t(parents), dcode.data(), t.img.data());
  QMBED_HIP(hipGetLastError());
  QMBED_HIP(hipStreamSynchronize(s));
  return true;
}

// Autotuner: H·v mode (orbit walk + sorted-table search, image table or
// compiled index stream) and lanes per row. Small sectors leave most of the
// 256 CUs idle with one lane per row; splitting a row's flip groups over 2-8
// lanes (reduced with cross-lane shuffles) restores parallelism. Each
// candidate FP64 kernel is timed with HIP events (the best of three rounds)
// and the fastest is kept only when it beats the default (the index stream
// with one lane per row, else the walk) by more than 10 % and 4 µs per
// product: the candidates sum a row's terms in different orders, so a pick
// made on timing noise (a sector whose products all take a few µs of launch
// latency) would change results in the last bits from run to run. An unused
// image table is released.
static void tune_lanes(Operator& op) {
  const SpinBasis1D& b = op.basis();
  const size_t n = b.len();
  if (n == 0) return;
  hipStream_t s;
  QMBED_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  // QMBED_HV_MODE = walk | image | stream | auto (default) restricts the
  // candidates
  const char* mode_env = std::getenv("QMBED_HV_MODE");
  const std::string mode = mode_env ? mode_env : "auto";
  // general groups, reversed rows and 64-bit states run on the compiled
  // stream only
  const bool stream_only = b.general() || b.reversed() || b.wide();
  const bool want_walk = (mode == "auto" || mode == "walk") && !stream_only;
  const bool have_img = (mode == "auto" || mode == "image") && build_image_table(op, s);
  const bool have_stream = (mode == "auto" || mode == "stream" || stream_only) && build_stream(op, s);
  if (stream_only && !have_stream) {
    (void)hipStreamDestroy(s);
    throw unsupported_backend(b.wide()      ? "the wide sector does not fit the compiled index stream"
                              : b.general() ? "the general symmetry sector does not fit the compiled index stream"
                                            : "the reversed sector does not fit the compiled index stream");
  }
  DeviceArray<double> x(2 * n), y(2 * n);
  QMBED_HIP(hipMemsetAsync(x.data(), 0, 2 * n * 8, s));
  hipEvent_t e0, e1;
  QMBED_HIP(hipEventCreate(&e0));
  QMBED_HIP(hipEventCreate(&e1));
  const int sym = kernel_sym(b);
  const bool sh = phased(b);
  float best = 1e30f, dflt = 1e30f;
  int best_lpr = 1, best_mode = 0;  // 0 walk, 1 image, 2 stream
  int dflt_mode = -1;
  for (int md : {2, 0, 1}) {
    if ((md == 0 && !want_walk) || (md == 1 && !have_img) || (md == 2 && !have_stream)) continue;
    for (int lpr : {1, 2, 4, 8}) {
      if (md == 2 && lpr > 1) break;  // the stream kernel runs one lane per row
      op.tables_->lpr = lpr;
      op.tables_->use_img = md == 1;
      op.tables_->use_stream = md == 2;
      HvArgs a = op.tables_->args(b);
      auto launch = [&]() {
        if (op.is_real()) launch_hv<double, 0, 1>(a, sym, sh, x.data(), y.data(), s);
        else launch_hv<double, 2, 1>(a, sym, sh, x.data(), y.data(), s);
      };
      launch();
      float ms = 1e30f;  // per product, best of three rounds of three
      for (int round = 0; round < 3; ++round) {
        QMBED_HIP(hipEventRecord(e0, s));
        for (int it = 0; it < 3; ++it) launch();
        QMBED_HIP(hipEventRecord(e1, s));
        QMBED_HIP(hipEventSynchronize(e1));
        float t = 0;
        QMBED_HIP(hipEventElapsedTime(&t, e0, e1));
        ms = std::min(ms, t / 3.0f);
      }
      if (dflt_mode < 0) {
        dflt_mode = md;  // the first candidate: one lane per row, stream or walk
        dflt = ms;
      }
      if (ms < best) {
        best = ms;
        best_lpr = lpr;
        best_mode = md;
      }
    }
  }
  if (best == 1e30f) throw unsupported_backend("QMBED_HV_MODE names an H·v mode this sector cannot use");
  if (dflt - best <= std::max(0.1f * dflt, 0.004f)) {
    best_mode = dflt_mode;
    best_lpr = 1;
  }
  op.tables_->lpr = best_lpr;
  op.tables_->use_img = best_mode == 1;
  op.tables_->use_stream = best_mode == 2;
  if (best_mode != 1) op.tables_->img.release();
  if (best_mode != 2) {
    op.tables_->stream.release();
    op.tables_->slice_ptr.release();
    op.tables_->aux.release();
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  QMBED_HIP(hipStreamSynchronize(s));
  (void)hipStreamDestroy(s);
}

bool Operator::uses_image_table() const { return tables_->use_img; }

bool Operator::uses_index_stream() const { return tables_->use_stream; }

int Operator::lanes_per_row() const { return tables_->lpr; }

static void scatter_apply(const Operator& op, const cplx* x, cplx* y, hipStream_t s) {
  const SpinBasis1D& b = op.basis();
  HvArgs a = op.tables_->args(b);
  if (a.n == 0) return;
  QMBED_HIP(hipMemsetAsync(y, 0, size_t(a.n) * sizeof(cplx), s));
  QMBED_WIDTH_LAUNCH(b, scatter_kernel, dim3((a.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a,
                     op.tables_->famp.data(), sym_code(b), phased(b) ? 1 : 0,
                     reinterpret_cast<const double2*>(x), reinterpret_cast<double*>(y));
  QMBED_HIP(hipGetLastError());
}

// The stored-CSR action when the operator uses a stored format; returns
// false (nothing launched) otherwise. cm: 0 real vectors, 1 complex
// (interleaved) vectors; f32: FP32 input (and FP32 output unless mixed).
static bool csr_apply(const Operator* op, int cm, bool f32, const void* x, void* y, hipStream_t s, const int* gate,
                      bool mixed = false) {
  auto& t = *op->tables_;
  if (!t.use_csr) return false;
  const_cast<Operator*>(op)->refresh_stored(s);
  const uint32_t n = uint32_t(op->dim());
  if (n == 0) return true;
  const dim3 grid(unsigned((size_t(n) * kCsrLanes + kBlock - 1) / kBlock)), block(kBlock);
  const int mode = cm == 0 ? 0 : (t.csr_complex ? 2 : 1);
  if (cm == 0 && t.csr_complex) throw unsupported_backend("real action requires an operator with real-valued storage");
  const uint64_t* p = t.csr_ptr.data();
  const uint32_t* c = t.csr_col.data();
  const double* v = t.csr_val.data();
#define QMBED_CSR(S, M, SX)                                                                                  \
  hipLaunchKernelGGL((csr_kernel<S, M, SX>), grid, block, 0, s, p, c, v, n, static_cast<const SX*>(x), \
                     static_cast<S*>(y), gate)
  if (!f32) {
    if (mode == 0) QMBED_CSR(double, 0, double);
    else if (mode == 1) QMBED_CSR(double, 1, double);
    else QMBED_CSR(double, 2, double);
  } else if (mixed) {
    if (mode == 0) QMBED_CSR(double, 0, float);
    else if (mode == 1) QMBED_CSR(double, 1, float);
    else QMBED_CSR(double, 2, float);
  } else {
    if (mode == 0) QMBED_CSR(float, 0, float);
    else if (mode == 1) QMBED_CSR(float, 1, float);
    else QMBED_CSR(float, 2, float);
  }
#undef QMBED_CSR
  QMBED_HIP(hipGetLastError());
  return true;
}

void Operator::apply(const double* x, double* y, hipStream_t s, const int* gate) const {
  if (!real_) throw unsupported_backend("real action requires an operator with real-valued storage");
  if (csr_apply(this, 0, false, x, y, s, gate)) return;
  HvArgs a = tables_->args(*basis_);
  a.gate = gate;
  launch_hv<double, 0, 1>(a, kernel_sym(*basis_), phased(*basis_), x, y, s);
}

void Operator::apply(const float* x, float* y, hipStream_t s, const int* gate) const {
  if (!real_) throw unsupported_backend("real action requires an operator with real-valued storage");
  if (csr_apply(this, 0, true, x, y, s, gate)) return;
  HvArgs a = tables_->args(*basis_);
  a.gate = gate;
  launch_hv<float, 0, 1>(a, kernel_sym(*basis_), phased(*basis_), x, y, s);
}

void Operator::apply(const cplx* x, cplx* y, hipStream_t s, const int* gate) const {
  if (csr_apply(this, 1, false, x, y, s, gate)) return;
  if (!tables_->hermitian) {
    scatter_apply(*this, x, y, s);
    return;
  }
  HvArgs a = tables_->args(*basis_);
  a.gate = gate;
  const double* xd = reinterpret_cast<const double*>(x);
  double* yd = reinterpret_cast<double*>(y);
  if (real_) launch_hv<double, 1, 1>(a, kernel_sym(*basis_), phased(*basis_), xd, yd, s);
  else launch_hv<double, 2, 1>(a, kernel_sym(*basis_), phased(*basis_), xd, yd, s);
}

void Operator::apply(const std::complex<float>* x, std::complex<float>* y, hipStream_t s, const int* gate) const {
  if (csr_apply(this, 1, true, x, y, s, gate)) return;
  if (!tables_->hermitian)
    throw unsupported_backend("single-precision action of a non-Hermitian sector matrix is not provided");
  HvArgs a = tables_->args(*basis_);
  a.gate = gate;
  const float* xf = reinterpret_cast<const float*>(x);
  float* yf = reinterpret_cast<float*>(y);
  if (real_) launch_hv<float, 1, 1>(a, kernel_sym(*basis_), phased(*basis_), xf, yf, s);
  else launch_hv<float, 2, 1>(a, kernel_sym(*basis_), phased(*basis_), xf, yf, s);
}

bool Operator::apply_mixed(const float* x, double* y, bool complex_vectors, hipStream_t s, const int* gate) const {
  if (tables_->use_csr) {
    if (!complex_vectors && !real_) throw unsupported_backend("real action requires an operator with real-valued storage");
    return csr_apply(this, complex_vectors ? 1 : 0, true, x, y, s, gate, true);
  }
  if (!tables_->use_stream || !tables_->hermitian) return false;
  HvArgs a = tables_->args(*basis_);
  a.gate = gate;
  const int sym = kernel_sym(*basis_);
  const bool sh = phased(*basis_);
  if (a.n == 0) return true;
  if (!complex_vectors) {
    if (!real_) throw unsupported_backend("real action requires an operator with real-valued storage");
    launch_hv_stream<double, 0, 1, float>(a, sym, sh, x, y, s);
  } else if (real_) {
    launch_hv_stream<double, 1, 1, float>(a, sym, sh, x, y, s);
  } else {
    launch_hv_stream<double, 2, 1, float>(a, sym, sh, x, y, s);
  }
  return true;
}

bool Operator::apply_batch_interleaved(const cplx* x, cplx* y, size_t ld, int nvec, hipStream_t s) const {
  if (!tables_->use_stream_interleaved || !tables_->hermitian) return false;
  HvArgs a = tables_->args(*basis_);
  a.stream = tables_->stream.data();
  a.slice_ptr = tables_->slice_ptr.data();
  a.img = nullptr;
  a.ld = ld;
  if (a.n == 0) return true;
  const int sym = kernel_sym(*basis_);
  const bool sh = phased(*basis_);
  const double* xd = reinterpret_cast<const double*>(x);
  double* yd = reinterpret_cast<double*>(y);
  int done = 0;
  while (done < nvec) {
    const int left = nvec - done;
    const double* xs = xd + 2 * size_t(done);
    double* ys = yd + 2 * size_t(done);
    // eight columns of complex FP64 fill one 128-byte line per gathered
    // target row; the remainder goes by four and one
    if (left >= 8) {
      if (real_) launch_hv_stream<double, 1, 8, double, true>(a, sym, sh, xs, ys, s);
      else launch_hv_stream<double, 2, 8, double, true>(a, sym, sh, xs, ys, s);
      done += 8;
    } else if (left >= 4) {
      if (real_) launch_hv_stream<double, 1, 4, double, true>(a, sym, sh, xs, ys, s);
      else launch_hv_stream<double, 2, 4, double, true>(a, sym, sh, xs, ys, s);
      done += 4;
    } else {
      if (real_) launch_hv_stream<double, 1, 1, double, true>(a, sym, sh, xs, ys, s);
      else launch_hv_stream<double, 2, 1, double, true>(a, sym, sh, xs, ys, s);
      done += 1;
    }
  }
  return true;
}

bool Operator::scan_batch_supported(hipStream_t s) const {
  auto& t = *tables_;
  if (!parameterized() || dim() == 0 || !parts_real_ || !basis_->real_characters() || phased(*basis_) ||
      basis_->general())
    return false;
  for (bool h : part_hermitian_)
    if (!h) return false;
  if (t.diag_parts != int(components_.size()) + 1) return false;
  if (t.stream.size() == 0 && !build_stream(const_cast<Operator&>(*this), s)) return false;
  // LDS of an 8-lane launch: the lane tables and the two small tables
  return t.compact && t.aux.size() == 0 && sizeof(double) * (8 * size_t(t.n_amp) + 81) <= 48 * 1024;
}

int Operator::scan_batch_max_width() const {
  return sizeof(double) * (16 * size_t(tables_->n_amp) + 81) <= 48 * 1024 ? 16 : 8;
}

int Operator::scan_batch_amp_count() const { return tables_->n_amp; }

std::vector<double> Operator::scan_amp_table(const std::vector<cplx>& c) const {
  auto& t = *tables_;
  const int na = t.n_amp;
  if (c.size() != components_.size())
    throw dimension_mismatch("received " + std::to_string(c.size()) + " operator coefficients for " +
                             std::to_string(components_.size()) + " components");
  for (const auto& v : c)
    if (!std::isfinite(v.real()) || v.imag() != 0.0) throw invalid_options("batched scan coefficients must be finite and real");
  // the gather table in upload order (diagonal groups, then off-diagonal)
  std::vector<double> amps(static_cast<size_t>(na));
  size_t k = 0;
  for (auto* v : {&diag_, &off_})
    for (auto& g : *v)
      for (size_t pat = 0; pat < g.parts[0].size(); ++pat) {
        cplx a = g.parts[0][pat];
        for (size_t p = 0; p < c.size(); ++p) a += c[p] * g.parts[p + 1][pat];
        amps[k++] = a.real();
      }
  if (k != size_t(na)) throw internal_state("batched scan amplitude table of the wrong length");
  return amps;
}

void Operator::scan_point_diagonal(const std::vector<cplx>& c, double* d_diag, hipStream_t s) const {
  auto& t = *tables_;
  const size_t n = dim();
  if (c.size() != components_.size())
    throw dimension_mismatch("received " + std::to_string(c.size()) + " operator coefficients for " +
                             std::to_string(components_.size()) + " components");
  DiagCoefs kc;
  kc.parts = t.diag_parts;
  kc.real_parts = 1;
  kc.complex_out = 0;
  kc.re[0] = 1.0;
  kc.im[0] = 0.0;
  for (size_t p = 0; p < c.size(); ++p) {
    kc.re[p + 1] = c[p].real();
    kc.im[p + 1] = 0.0;
  }
  const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(diag_combine_kernel, dim3(unsigned(blocks)), dim3(kBlock), 0, s, t.part_diag.data(), n, kc,
                     d_diag);
  QMBED_HIP(hipGetLastError());
}

void Operator::scan_batch_tables(const std::vector<std::vector<cplx>>& coefs, double* d_amps, double* d_diags,
                                 hipStream_t s) const {
  auto& t = *tables_;
  const size_t n = dim();
  const int na = t.n_amp;
  std::vector<double> amps(coefs.size() * size_t(na));
  for (size_t b = 0; b < coefs.size(); ++b) {
    const auto a = scan_amp_table(coefs[b]);
    std::copy(a.begin(), a.end(), amps.begin() + long(b * size_t(na)));
    scan_point_diagonal(coefs[b], d_diags + b * 2 * n, s);
  }
  // staged through the tables' pinned area (free once its last copy has
  // run), so the copy is asynchronous
  if (t.staged) QMBED_HIP(hipEventSynchronize(t.staged));
  t.stage.ensure(amps.size());
  std::copy(amps.begin(), amps.end(), t.stage.p);
  QMBED_HIP(hipMemcpyAsync(d_amps, t.stage.p, amps.size() * sizeof(double), hipMemcpyHostToDevice, s));
  if (!t.staged) QMBED_HIP(hipEventCreateWithFlags(&t.staged, hipEventDisableTiming));
  QMBED_HIP(hipEventRecord(t.staged, s));
}

// LDS bytes of an on-the-fly scan launch of `width` lanes: the lanes'
// off-diagonal amplitude tables, sqrt(i) and 1/sqrt(i), the off-diagonal
// group words and the directory cells.
static size_t scan_walk_smem(const Operator& op, int width, int off_base) {
  const HvArgs a = op.tables_->args(op.basis());
  return sizeof(double) * (size_t(width) * size_t(a.n_amp - off_base) + 130) +
         sizeof(uint32_t) * (4 * size_t(a.n_off) + 2 * size_t(a.ncells));
}

static int diagonal_amp_count(const Operator& op) {
  int off_base = 0;
  for (auto& grp : op.diagonal_groups()) off_base += int(grp.amp.size());
  return off_base;
}

// Whether the batched scan action runs on the fly (orbit walk + state
// directory per target) instead of reading the compiled stream: asked for
// with QMBED_SCAN_HV=walk, for sectors of built-in symmetries, ascending
// rows and 32-bit states without character phases whose 16-lane launch fits
// the LDS (otherwise the stream form runs; both give the same results).
static bool scan_on_the_fly(const Operator& op) {
  auto& t = *op.tables_;
  if (t.scan_walk < 0) {
    const char* m = std::getenv("QMBED_SCAN_HV");
    const SpinBasis1D& b = op.basis();
    t.scan_walk = m && std::string(m) == "walk" && !b.general() && !b.reversed() && !b.wide() && !phased(b) &&
                          scan_walk_smem(op, 16, diagonal_amp_count(op)) <= 64 * 1024
                      ? 1
                      : 0;
  }
  return t.scan_walk == 1;
}

static void launch_scan_walk(const Operator& op, const void* x, bool fp32, int width, const ScanBatch& sb,
                             hipStream_t s) {
  auto& t = *op.tables_;
  const SpinBasis1D& b = op.basis();
  HvArgs a = t.args(b);
  a.img = nullptr;
  a.stream = nullptr;
  if (a.n == 0) return;
  if (t.size_rsq.size() == 0) {
    std::vector<double> h(65, 0.0);
    for (int i = 1; i < 65; ++i) h[size_t(i)] = 1.0 / std::sqrt(double(i));
    t.size_rsq.resize(65);
    QMBED_HIP(hipMemcpy(t.size_rsq.data(), h.data(), 65 * 8, hipMemcpyHostToDevice));
  }
  const int off_base = diagonal_amp_count(op);
  const size_t sm = scan_walk_smem(op, width, off_base);
  if (sm > 64 * 1024) throw internal_state("on-the-fly scan action beyond the LDS of its tables");
  // the kernel's lanes per row
  const size_t lanes = size_t(a.n) * size_t(scan_walk_lanes(width));
  const dim3 grid(unsigned((lanes + kBlock - 1) / kBlock)), block(kBlock);
  const int sym = kernel_sym(b);
  char kname[64];
  std::snprintf(kname, sizeof kname, "hv_scan_walk_kernel<%s, %d, %d>", fp32 ? "float" : "double", width,
                sym < 0 || sym > 3 ? 3 : sym);
  const int timed = launch_timer_begin(kname, s);
  const double* rsq = t.size_rsq.data();
#define QMBED_WALK_LAUNCH(SX, W)                                                                                \
  switch (sym) {                                                                                               \
    case 0: hipLaunchKernelGGL((hv_scan_walk_kernel<SX, W, 0>), grid, block, sm, s, a, sb, static_cast<const SX*>(x), off_base, rsq); break; \
    case 1: hipLaunchKernelGGL((hv_scan_walk_kernel<SX, W, 1>), grid, block, sm, s, a, sb, static_cast<const SX*>(x), off_base, rsq); break; \
    case 2: hipLaunchKernelGGL((hv_scan_walk_kernel<SX, W, 2>), grid, block, sm, s, a, sb, static_cast<const SX*>(x), off_base, rsq); break; \
    default: hipLaunchKernelGGL((hv_scan_walk_kernel<SX, W, 3>), grid, block, sm, s, a, sb, static_cast<const SX*>(x), off_base, rsq); break; \
  }
  if (fp32) {
    switch (width) {
      case 1: QMBED_WALK_LAUNCH(float, 1) break;
      case 2: QMBED_WALK_LAUNCH(float, 2) break;
      case 4: QMBED_WALK_LAUNCH(float, 4) break;
      case 8: QMBED_WALK_LAUNCH(float, 8) break;
      default: QMBED_WALK_LAUNCH(float, 16) break;
    }
  } else {
    switch (width) {
      case 1: QMBED_WALK_LAUNCH(double, 1) break;
      case 2: QMBED_WALK_LAUNCH(double, 2) break;
      case 4: QMBED_WALK_LAUNCH(double, 4) break;
      case 8: QMBED_WALK_LAUNCH(double, 8) break;
      default: QMBED_WALK_LAUNCH(double, 16) break;
    }
  }
#undef QMBED_WALK_LAUNCH
  QMBED_HIP(hipGetLastError());
  launch_timer_end(timed, s);
}

void Operator::apply_scan_batch(const void* x, bool fp32, int width, const double* const* amps,
                                const double* const* diags, double* const* y, const int* const* gates,
                                hipStream_t s) const {
  auto& t = *tables_;
  if (!(width == 1 || width == 2 || width == 4 || width == 8 || width == 16) || width > kMaxScanBatch)
    throw internal_state("batched scan width must be 1, 2, 4, 8 or 16");
  if (width > scan_batch_max_width()) throw internal_state("batched scan width beyond the LDS of its tables");
  if (scan_on_the_fly(*this)) {
    ScanBatch sb{};
    for (int b = 0; b < width; ++b) {
      sb.amp[b] = amps[b];
      sb.diag[b] = diags[b];
      sb.y[b] = y[b];
      sb.gate[b] = gates ? gates[b] : nullptr;
    }
    launch_scan_walk(*this, x, fp32, width, sb, s);
    return;
  }
  if (!t.compact || t.stream.size() == 0) throw internal_state("batched scan action without a compact stream");
  HvArgs a = t.args(*basis_);
  a.stream = t.stream.data();
  a.slice_ptr = t.slice_ptr.data();
  a.img = nullptr;
  if (a.n == 0) return;
  ScanBatch sb{};
  for (int b = 0; b < width; ++b) {
    sb.amp[b] = amps[b];
    sb.diag[b] = diags[b];
    sb.y[b] = y[b];
    sb.gate[b] = gates ? gates[b] : nullptr;
  }
  const size_t n_slices = (size_t(a.n) + kSlice - 1) / kSlice;
  const dim3 grid(unsigned((n_slices * kSlice + kBlock - 1) / kBlock)), block(kBlock);
  const size_t sm = sizeof(double) * (size_t(width) * size_t(a.n_amp) + 81);
  const int sym = kernel_sym(*basis_);
  char kname[64];
  std::snprintf(kname, sizeof kname, "hv_scan_batch_kernel<%s, %d, %d>", fp32 ? "float" : "double", width,
                sym < 0 || sym > 3 ? 3 : sym);
  const int timed = launch_timer_begin(kname, s);
#define QMBED_SCAN_LAUNCH(SX, W)                                                                                   \
  switch (sym) {                                                                                                  \
    case 0: hipLaunchKernelGGL((hv_scan_batch_kernel<SX, W, 0>), grid, block, sm, s, a, sb, static_cast<const SX*>(x)); break; \
    case 1: hipLaunchKernelGGL((hv_scan_batch_kernel<SX, W, 1>), grid, block, sm, s, a, sb, static_cast<const SX*>(x)); break; \
    case 2: hipLaunchKernelGGL((hv_scan_batch_kernel<SX, W, 2>), grid, block, sm, s, a, sb, static_cast<const SX*>(x)); break; \
    default: hipLaunchKernelGGL((hv_scan_batch_kernel<SX, W, 3>), grid, block, sm, s, a, sb, static_cast<const SX*>(x)); break; \
  }
  if (fp32) {
    switch (width) {
      case 1: QMBED_SCAN_LAUNCH(float, 1) break;
      case 2: QMBED_SCAN_LAUNCH(float, 2) break;
      case 4: QMBED_SCAN_LAUNCH(float, 4) break;
      case 8: QMBED_SCAN_LAUNCH(float, 8) break;
      default: QMBED_SCAN_LAUNCH(float, 16) break;
    }
  } else {
    switch (width) {
      case 1: QMBED_SCAN_LAUNCH(double, 1) break;
      case 2: QMBED_SCAN_LAUNCH(double, 2) break;
      case 4: QMBED_SCAN_LAUNCH(double, 4) break;
      case 8: QMBED_SCAN_LAUNCH(double, 8) break;
      default: QMBED_SCAN_LAUNCH(double, 16) break;
    }
  }
#undef QMBED_SCAN_LAUNCH
  QMBED_HIP(hipGetLastError());
  launch_timer_end(timed, s);
}

void Operator::apply_scan_mixed(const float* x32, int w32, const double* x64, int w64, const double* const* amps,
                                const double* const* diags, double* const* y, const int* const* gates,
                                hipStream_t s) const {
  auto& t = *tables_;
  auto pow2_or_zero = [](int w, int cap) { return w == 0 || ((w & (w - 1)) == 0 && w <= cap); };
  if (!pow2_or_zero(w32, 16) || !pow2_or_zero(w64, 8) || w32 + w64 == 0)
    throw internal_state("mixed scan widths must be 0, 1, 2, 4, 8 or 16 (FP32) and 0, 1, 2, 4 or 8 (FP64)");
  if (w64 == 0) return apply_scan_batch(x32, true, w32, amps, diags, y, gates, s);
  if (w32 == 0) return apply_scan_batch(x64, false, w64, amps, diags, y, gates, s);
  if (scan_on_the_fly(*this)) {
    // one launch per precision (each lane's terms and order are those of
    // the mixed launch)
    apply_scan_batch(x32, true, w32, amps, diags, y, gates, s);
    apply_scan_batch(x64, false, w64, amps + w32, diags + w32, y + w32, gates ? gates + w32 : nullptr, s);
    return;
  }
  const int B = w32 + w64;
  if (B > kMaxScanBatch) throw internal_state("mixed scan launch of more lanes than a batch holds");
  if (sizeof(double) * (size_t(B) * size_t(t.n_amp) + 81) > 64 * 1024)
    throw internal_state("mixed scan width beyond the LDS of its tables");
  if (!t.compact || t.stream.size() == 0) throw internal_state("batched scan action without a compact stream");
  HvArgs a = t.args(*basis_);
  a.stream = t.stream.data();
  a.slice_ptr = t.slice_ptr.data();
  a.img = nullptr;
  if (a.n == 0) return;
  ScanBatch sb{};
  for (int b = 0; b < B; ++b) {
    sb.amp[b] = amps[b];
    sb.diag[b] = diags[b];
    sb.y[b] = y[b];
    sb.gate[b] = gates ? gates[b] : nullptr;
  }
  const size_t n_slices = (size_t(a.n) + kSlice - 1) / kSlice;
  const dim3 grid(unsigned((n_slices * kSlice + kBlock - 1) / kBlock)), block(kBlock);
  const size_t sm = sizeof(double) * (size_t(B) * size_t(t.n_amp) + 81);
  const int sym = kernel_sym(*basis_);
  char kname[64];
  std::snprintf(kname, sizeof kname, "hv_scan_mixed_kernel<%d, %d, %d>", w32, w64, sym < 0 || sym > 3 ? 3 : sym);
  const int timed = launch_timer_begin(kname, s);
#define QMBED_MIXED_SYM(W32, W64)                                                                                    \
  switch (sym) {                                                                                                      \
    case 0: hipLaunchKernelGGL((hv_scan_mixed_kernel<W32, W64, 0>), grid, block, sm, s, a, sb, x32, x64); break;    \
    case 1: hipLaunchKernelGGL((hv_scan_mixed_kernel<W32, W64, 1>), grid, block, sm, s, a, sb, x32, x64); break;    \
    case 2: hipLaunchKernelGGL((hv_scan_mixed_kernel<W32, W64, 2>), grid, block, sm, s, a, sb, x32, x64); break;    \
    default: hipLaunchKernelGGL((hv_scan_mixed_kernel<W32, W64, 3>), grid, block, sm, s, a, sb, x32, x64); break;   \
  }
#define QMBED_MIXED_W64(W32)                 \
  switch (w64) {                             \
    case 1: QMBED_MIXED_SYM(W32, 1) break;   \
    case 2: QMBED_MIXED_SYM(W32, 2) break;   \
    case 4: QMBED_MIXED_SYM(W32, 4) break;   \
    default: QMBED_MIXED_SYM(W32, 8) break;  \
  }
  switch (w32) {
    case 1: QMBED_MIXED_W64(1) break;
    case 2: QMBED_MIXED_W64(2) break;
    case 4: QMBED_MIXED_W64(4) break;
    case 8: QMBED_MIXED_W64(8) break;
    default: QMBED_MIXED_W64(16) break;
  }
#undef QMBED_MIXED_W64
#undef QMBED_MIXED_SYM
  QMBED_HIP(hipGetLastError());
  launch_timer_end(timed, s);
}

bool prepare_interleaved_action(const Operator& op, hipStream_t s) {
  auto& t = *op.tables_;
  if (!t.hermitian || op.dim() == 0) return false;
  if (!t.use_stream_interleaved) {
    if (t.stream.size() == 0 && !build_stream(const_cast<Operator&>(op), s)) return false;
    t.use_stream_interleaved = true;
  }
  return true;
}

void Operator::apply_batch(const cplx* x, cplx* y, int nvec, hipStream_t s) const {
  if (tables_->use_csr) {
    for (int v = 0; v < nvec; ++v) apply(x + size_t(v) * dim(), y + size_t(v) * dim(), s);
    return;
  }
  if (!tables_->hermitian) {
    for (int v = 0; v < nvec; ++v) scatter_apply(*this, x + size_t(v) * dim(), y + size_t(v) * dim(), s);
    return;
  }
  HvArgs a = tables_->args(*basis_);
  const double* xd = reinterpret_cast<const double*>(x);
  double* yd = reinterpret_cast<double*>(y);
  const int sym = kernel_sym(*basis_);
  const bool sh = phased(*basis_);
  int done = 0;
  while (done < nvec) {
    int left = nvec - done;
    const double* xs = xd + size_t(2) * done * dim();
    double* ys = yd + size_t(2) * done * dim();
    if (left >= 4) {
      if (real_) launch_hv<double, 1, 4>(a, sym, sh, xs, ys, s);
      else launch_hv<double, 2, 4>(a, sym, sh, xs, ys, s);
      done += 4;
    } else if (left >= 2) {
      if (real_) launch_hv<double, 1, 2>(a, sym, sh, xs, ys, s);
      else launch_hv<double, 2, 2>(a, sym, sh, xs, ys, s);
      done += 2;
    } else {
      if (real_) launch_hv<double, 1, 1>(a, sym, sh, xs, ys, s);
      else launch_hv<double, 2, 1>(a, sym, sh, xs, ys, s);
      done += 1;
    }
  }
}

static void run_columns(const Operator& op, cplx* dense, double* colsum, double* dre, double* dim_, cplx shift,
                        hipStream_t s, uint32_t* count = nullptr) {
  const SpinBasis1D& b = op.basis();
  HvArgs a = op.tables_->args(b);
  if (a.n == 0) return;
  DeviceArray<int> overflow(1);
  QMBED_HIP(hipMemsetAsync(overflow.data(), 0, sizeof(int), s));
  dim3 grid((a.n + kBlock - 1) / kBlock);
  QMBED_WIDTH_LAUNCH(b, column_entries_kernel, grid, dim3(kBlock), 0, s, a, op.tables_->famp.data(), sym_code(b),
                     phased(b) ? 1 : 0, dense, colsum, dre, dim_, overflow.data(), shift, count);
  QMBED_HIP(hipGetLastError());
  int of = 0;
  QMBED_HIP(hipMemcpyAsync(&of, overflow.data(), sizeof(int), hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  if (of) throw unsupported_backend("a column has more distinct entries than the column-entry kernel holds");
}

size_t Operator::stored_entries(hipStream_t s) const {
  size_t n = dim();
  if (n == 0) return 0;
  DeviceArray<uint32_t> cnt(n);
  run_columns(*this, nullptr, nullptr, nullptr, nullptr, cplx(0, 0), s, cnt.data());
  std::vector<uint32_t> h(n);
  QMBED_HIP(hipMemcpy(h.data(), cnt.data(), n * 4, hipMemcpyDeviceToHost));
  size_t tot = 0;
  for (uint32_t c : h) tot += c;
  return tot;
}

// Assemble the stored sector matrix as CSR rows (OperatorBuilder::build
// with a stored format, src/operator/mod.rs:2091-2216): count and fill the
// CSC columns, then read them as rows, conjugated, when the matrix is
// Hermitian, or order the entries by row (a stable radix sort keeps the
// column order within a row) when it is not.
void Operator::refresh_stored(hipStream_t s) {
  auto& t = *tables_;
  if (t.csr_version == t.version) return;
  const SpinBasis1D& b = *basis_;
  const size_t n = b.len();
  t.csr_ptr.resize(n + 1);
  if (n == 0) {
    const uint64_t z = 0;
    QMBED_HIP(hipMemcpy(t.csr_ptr.data(), &z, 8, hipMemcpyHostToDevice));
    t.csr_nnz = 0;
    t.csr_version = t.version;
    return;
  }
  HvArgs a = t.args(b);
  DeviceArray<uint32_t> cnt(n);
  DeviceArray<int> overflow(1);
  QMBED_HIP(hipMemsetAsync(overflow.data(), 0, sizeof(int), s));
  const dim3 grid(unsigned((n + kBlock - 1) / kBlock));
  QMBED_WIDTH_LAUNCH(b, column_entries_kernel, grid, dim3(kBlock), 0, s, a, t.famp.data(), sym_code(b),
                     phased(b) ? 1 : 0, nullptr, nullptr, nullptr, nullptr, overflow.data(), cplx(0, 0), cnt.data());
  QMBED_HIP(hipGetLastError());
  std::vector<uint32_t> h(n);
  QMBED_HIP(hipMemcpyAsync(h.data(), cnt.data(), n * 4, hipMemcpyDeviceToHost, s));
  int of = 0;
  QMBED_HIP(hipMemcpyAsync(&of, overflow.data(), sizeof(int), hipMemcpyDeviceToHost, s));
  QMBED_HIP(hipStreamSynchronize(s));
  if (of) throw unsupported_backend("a column has more distinct entries than the column-entry kernel holds");
  std::vector<uint64_t> ptr(n + 1, 0);
  for (size_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + h[i];
  const uint64_t nnz = ptr[n];
  if (nnz >= (uint64_t(1) << 32)) throw unsupported_backend("the stored matrix exceeds 2^32 entries");
  DeviceArray<uint64_t> cptr(n + 1);
  QMBED_HIP(hipMemcpyAsync(cptr.data(), ptr.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
  DeviceArray<uint32_t> rows(std::max<uint64_t>(nnz, 1));
  DeviceArray<double2> vals(std::max<uint64_t>(nnz, 1));
  QMBED_WIDTH_LAUNCH(b, column_fill_kernel, grid, dim3(kBlock), 0, s, a, t.famp.data(), sym_code(b), phased(b) ? 1 : 0,
                     cptr.data(), rows.data(), vals.data(), overflow.data());
  QMBED_HIP(hipGetLastError());
  t.csr_complex = !real_;
  t.csr_val.resize(std::max<uint64_t>(nnz, 1) * (t.csr_complex ? 2 : 1));
  const unsigned eg = unsigned((nnz + 255) / 256);
  if (t.hermitian) {
    QMBED_HIP(hipMemcpyAsync(t.csr_ptr.data(), cptr.data(), (n + 1) * 8, hipMemcpyDeviceToDevice, s));
    t.csr_col = std::move(rows);
    if (nnz)
      hipLaunchKernelGGL(csr_values_kernel, dim3(eg), dim3(256), 0, s, vals.data(), nullptr, nullptr, nnz,
                         t.csr_complex ? 1 : 0, t.csr_val.data(), nullptr);
  } else {
    DeviceArray<uint32_t> ccols(std::max<uint64_t>(nnz, 1)), perm_in(std::max<uint64_t>(nnz, 1)),
        perm(std::max<uint64_t>(nnz, 1)), keys(std::max<uint64_t>(nnz, 1));
    hipLaunchKernelGGL(csc_columns_kernel, grid, dim3(kBlock), 0, s, cptr.data(), uint32_t(n), ccols.data());
    std::vector<uint32_t> iota(nnz);
    std::iota(iota.begin(), iota.end(), 0u);
    QMBED_HIP(hipMemcpyAsync(perm_in.data(), iota.data(), nnz * 4, hipMemcpyHostToDevice, s));
    int bits = 1;
    while ((size_t(1) << bits) < n) ++bits;
    size_t tmp_bytes = 0;
    QMBED_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, rows.data(), keys.data(), perm_in.data(),
                                                 perm.data(), int(nnz), 0, bits, s));
    DeviceArray<uint8_t> tmp(std::max<size_t>(tmp_bytes, 1));
    QMBED_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.data(), tmp_bytes, rows.data(), keys.data(), perm_in.data(),
                                                 perm.data(), int(nnz), 0, bits, s));
    t.csr_col.resize(std::max<uint64_t>(nnz, 1));
    if (nnz)
      hipLaunchKernelGGL(csr_values_kernel, dim3(eg), dim3(256), 0, s, vals.data(), perm.data(), ccols.data(), nnz,
                         t.csr_complex ? 1 : 0, t.csr_val.data(), t.csr_col.data());
    DeviceArray<uint32_t> rc(n);
    QMBED_HIP(hipMemsetAsync(rc.data(), 0, n * 4, s));
    if (nnz) hipLaunchKernelGGL(row_count_kernel, dim3(eg), dim3(256), 0, s, keys.data(), nnz, rc.data());
    QMBED_HIP(hipMemcpyAsync(h.data(), rc.data(), n * 4, hipMemcpyDeviceToHost, s));
    QMBED_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + h[i];
    QMBED_HIP(hipMemcpyAsync(t.csr_ptr.data(), ptr.data(), (n + 1) * 8, hipMemcpyHostToDevice, s));
  }
  QMBED_HIP(hipGetLastError());
  QMBED_HIP(hipStreamSynchronize(s));
  t.csr_nnz = nnz;
  t.csr_version = t.version;
}

void Operator::use_format(MatrixFormat f, hipStream_t s) {
  format_ = f;
  auto& t = *tables_;
  t.use_csr = f != MatrixFormat::MatrixFree;
  if (!t.use_csr) {
    t.csr_ptr.release();
    t.csr_col.release();
    t.csr_val.release();
    t.csr_version = -1;
    return;
  }
  refresh_stored(s);
  if (f == MatrixFormat::Dia) {
    // the reference refuses a banded layout with more distinct diagonals
    // than 2·min(rows, columns) (src/operator/mod.rs:2174-2186)
    const size_t n = dim();
    std::vector<uint64_t> ptr(n + 1);
    std::vector<uint32_t> col(t.csr_nnz);
    QMBED_HIP(hipMemcpy(ptr.data(), t.csr_ptr.data(), (n + 1) * 8, hipMemcpyDeviceToHost));
    if (t.csr_nnz) QMBED_HIP(hipMemcpy(col.data(), t.csr_col.data(), t.csr_nnz * 4, hipMemcpyDeviceToHost));
    std::vector<char> seen(2 * n + 1, 0);
    size_t distinct = 0;
    for (size_t r = 0; r < n; ++r)
      for (uint64_t k = ptr[r]; k < ptr[r + 1]; ++k) {
        const size_t d = size_t(long(r) - long(col[k]) + long(n));
        if (!seen[d]) {
          seen[d] = 1;
          ++distinct;
        }
      }
    if (distinct > std::max<size_t>(2 * n, 1)) {
      use_format(MatrixFormat::MatrixFree, s);
      throw unsupported_backend("the requested operator is not usefully diagonal-banded");
    }
  }
}

std::vector<Operator::Triplet> Operator::triplets(hipStream_t s) const {
  auto* self = const_cast<Operator*>(this);
  const bool had = tables_->use_csr;
  tables_->use_csr = true;
  self->refresh_stored(s);
  const size_t n = dim();
  auto& t = *tables_;
  std::vector<uint64_t> ptr(n + 1, 0);
  std::vector<uint32_t> col(t.csr_nnz);
  std::vector<double> val(t.csr_nnz * (t.csr_complex ? 2 : 1));
  QMBED_HIP(hipMemcpy(ptr.data(), t.csr_ptr.data(), (n + 1) * 8, hipMemcpyDeviceToHost));
  if (t.csr_nnz) {
    QMBED_HIP(hipMemcpy(col.data(), t.csr_col.data(), col.size() * 4, hipMemcpyDeviceToHost));
    QMBED_HIP(hipMemcpy(val.data(), t.csr_val.data(), val.size() * 8, hipMemcpyDeviceToHost));
  }
  tables_->use_csr = had;
  std::vector<Triplet> out;
  out.reserve(t.csr_nnz);
  for (size_t r = 0; r < n; ++r)
    for (uint64_t k = ptr[r]; k < ptr[r + 1]; ++k)
      out.push_back(Triplet{uint32_t(r), col[k],
                            t.csr_complex ? cplx(val[2 * k], val[2 * k + 1]) : cplx(val[k], 0.0)});
  return out;
}

static_assert(SectorBatchAction::kRows == kBlock, "batched blocks are H·v blocks");

SectorBatchAction::SectorBatchAction(const std::vector<Operator*>& ops, const std::vector<size_t>& offsets,
                                     hipStream_t s) {
  if (ops.size() != offsets.size()) throw dimension_mismatch("sector batch: one offset per operator is required");
  if (ops.size() >= (size_t(1) << 16)) throw unsupported_backend("sector batch: at most 65535 sectors");
  count_ = int(ops.size());
  std::vector<BatchSector> host(ops.size());
  for (size_t i = 0; i < ops.size(); ++i) {
    Operator& op = *ops[i];
    const SpinBasis1D& b = op.basis();
    if (!op.hermitian()) throw non_hermitian();
    const int sym = kernel_sym(b);
    if (i == 0) sym_ = sym;
    if (sym != sym_) throw unsupported_backend("sector batch: every sector must use the same symmetry kind");
    auto& t = *op.tables_;
    if (t.stream.size() == 0 && !build_stream(op, s))
      throw unsupported_backend("sector batch: a sector does not fit the compiled index stream");
    const uint32_t nb = uint32_t((b.len() + kRows - 1) / kRows);
    if (nb > 0xFFFFu) throw unsupported_backend("sector batch: a sector exceeds 65535 blocks");
    blocks_.push_back(nb);
    BatchSector bs{};
    bs.a = t.args(b);
    bs.a.stream = t.stream.data();
    bs.a.slice_ptr = t.slice_ptr.data();
    bs.a.img = nullptr;
    bs.a.gate = nullptr;
    bs.offset = offsets[i];
    bs.cmode = op.is_real() ? 1 : 2;
    bs.shift = phased(b) ? 1 : 0;
    host[i] = bs;
    const size_t G = size_t(bs.a.n_diag + bs.a.n_off);
    smem_ = std::max(smem_,
                     sizeof(double) * (2 * size_t(bs.a.n_amp) + 2 * kMaxChars + 65 + 16) + sizeof(uint32_t) * 4 * G);
  }
  sec_.resize(std::max<size_t>(1, host.size() * sizeof(BatchSector)));
  if (!host.empty())
    QMBED_HIP(hipMemcpy(sec_.data(), host.data(), host.size() * sizeof(BatchSector), hipMemcpyHostToDevice));
}

void SectorBatchAction::apply(const uint32_t* d_items, int nitems, const double* x, double* y, hipStream_t s) const {
  if (nitems <= 0) return;
  const BatchSector* sec = reinterpret_cast<const BatchSector*>(sec_.data());
  switch (sym_) {
    case 0: hipLaunchKernelGGL(hv_stream_batch_kernel<0>, dim3(nitems), dim3(kBlock), smem_, s, sec, d_items, x, y); break;
    case 1: hipLaunchKernelGGL(hv_stream_batch_kernel<1>, dim3(nitems), dim3(kBlock), smem_, s, sec, d_items, x, y); break;
    case 2: hipLaunchKernelGGL(hv_stream_batch_kernel<2>, dim3(nitems), dim3(kBlock), smem_, s, sec, d_items, x, y); break;
    default: hipLaunchKernelGGL(hv_stream_batch_kernel<3>, dim3(nitems), dim3(kBlock), smem_, s, sec, d_items, x, y); break;
  }
  QMBED_HIP(hipGetLastError());
}

void Operator::materialize_dense(cplx* dense, hipStream_t s) const {
  size_t n = dim();
  QMBED_HIP(hipMemsetAsync(dense, 0, n * n * sizeof(cplx), s));
  run_columns(*this, dense, nullptr, nullptr, nullptr, cplx(0, 0), s);
}

std::pair<cplx, double> Operator::shifted_trace_and_one_norm(hipStream_t s) const {
  size_t n = dim();
  if (n == 0) return {cplx(0, 0), 0.0};
  DeviceArray<double> dre(n), dimv(n), colsum(n);
  run_columns(*this, nullptr, nullptr, dre.data(), dimv.data(), cplx(0, 0), s);
  std::vector<double> hre(n), him(n);
  QMBED_HIP(hipMemcpy(hre.data(), dre.data(), n * 8, hipMemcpyDeviceToHost));
  QMBED_HIP(hipMemcpy(him.data(), dimv.data(), n * 8, hipMemcpyDeviceToHost));
  cplx trace(0, 0);
  for (size_t i = 0; i < n; ++i) {
    // diagonal entries with |v| <= eps are not stored by the reference
    cplx d(hre[i], him[i]);
    if (std::abs(d) > kEps) trace += d;
  }
  cplx shift = trace / double(n);
  run_columns(*this, nullptr, colsum.data(), nullptr, nullptr, shift, s);
  std::vector<double> cs(n);
  QMBED_HIP(hipMemcpy(cs.data(), colsum.data(), n * 8, hipMemcpyDeviceToHost));
  double m = 0;
  for (double v : cs) m = std::max(m, v);
  return {shift, m};
}

}  // namespace qmbed
