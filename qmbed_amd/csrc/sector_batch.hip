// Lockstep Lanczos over many symmetry sectors: the per-sector algorithm is
// the reference's fully reorthogonalized window (QMBED src/solve/mod.rs
// 1116-1456, DGKS criterion :1050-1073) with the early stop of eigsh.cpp;
// the batching is this library's. Every launch of a step covers all sectors
// still running, so sectors of ~10^5 states, each far too small to fill the
// GPU alone, share one grid per phase.
//
// Layout: the sectors' vectors are concatenated, each segment padded to a
// whole number of 2048-row tiles (zero rows stay zero through every phase),
// complex FP64 throughout (real sectors carry zero imaginary parts, which
// their real operators keep exactly zero). The Krylov basis is m such
// concatenated columns. A tile belongs to exactly one sector, so every
// per-sector reduction is a fixed-order sum of per-tile partials.
//
// One step, all running sectors at once (j is common to them):
//   H·v (sector-batched stream kernel) -> α partials -> α per sector ->
//   w₀ = w − α v_j − β v_{j−1} (+ ‖w₀‖² partials) -> overlaps V^H w₀ ->
//   per-sector sums -> w ← w₀ − V h (+ ‖w‖² partials) -> DGKS decision,
//   which lists the sectors needing a second pass -> second pass over the
//   listed sectors only (grid-stride launches whose work list the decision
//   kernel wrote; an empty list makes them no-ops) -> v_{j+1} = w/β.
// Coefficients reach the host through lagged read-backs; converged sectors
// leave the running set after their Ritz pairs pass the FP64 residual test.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <exception>
#include <numeric>
#include <thread>

#include "device_common.hpp"
#include "internal.hpp"

namespace qmbed {

std::vector<double> tridiagonal_eigvec(const std::vector<double>& alpha, const std::vector<double>& beta, double lam);
Eigensystem eigsh_on_device(const Operator& op, const EigshOptions& o, const double* d_initial, bool initial_complex,
                            const EigshControls& ctl, DeviceArray<double>& vectors, bool& complex_vectors,
                            hipStream_t s);

namespace {

constexpr int kT = 256;
constexpr int kRPT = 8;
constexpr int kTile = kT * kRPT;  // rows per tile
constexpr int kCT = 4;            // overlap columns per pass over a tile
constexpr int kMaxSectors = 4096;
constexpr int kMaxWindow = 1024;
constexpr int kSecondGrid = 1024;  // blocks of the second-pass launches
constexpr double kEps = 2.220446049250313e-16;

struct Seg {
  uint32_t tile0;   // first tile
  uint32_t ntiles;  // tiles of the segment
  uint32_t n;       // rows (dimension)
  uint32_t pad;
};

// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// a * b
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double nsq(double2 a) { return a.x * a.x + a.y * a.y; }

// sum over a wave, valid in lane 0 (shuffle tree: a fixed order)
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double2 wave_sum(double2 v) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o, 64);
    v.y += __shfl_down(v.y, o, 64);
  }
  return v;
}
// block sum, valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kT / 64; ++w) r += sh[w];
  __syncthreads();
  return r;
}

__device__ __forceinline__ size_t row_of(uint32_t tile, int i) {
  return size_t(tile) * kTile + size_t(i) * kT + threadIdx.x;
}

// ------------------------------------------------------------ vector ops ---

// column 0: the deterministic start of each sector (real part), zero padding
__global__ __launch_bounds__(kT) void kb_start(const uint32_t* __restrict__ tiles,
                                               const uint32_t* __restrict__ tile_sector, const Seg* __restrict__ seg,
                                               uint64_t seed, double2* __restrict__ v) {
  const uint32_t tile = tiles[blockIdx.x];
  const Seg g = seg[tile_sector[tile]];
  for (int i = 0; i < kRPT; ++i) {
    const size_t r = row_of(tile, i);
    const size_t local = r - size_t(g.tile0) * kTile;
    v[r] = make_double2(local < g.n ? deterministic_entry(local, seed) : 0.0, 0.0);
  }
}

// per-tile partials: MODE 0 Σ|x|², MODE 1 Re Σ conj(x) y, MODE 2 Σ|y − λ x|²
// with λ = lam[sector · stride + off]
template <int MODE>
__global__ __launch_bounds__(kT) void kb_tile_reduce(const uint32_t* __restrict__ tiles,
                                                     const uint32_t* __restrict__ tile_sector,
                                                     const double2* __restrict__ x, const double2* __restrict__ y,
                                                     const double* __restrict__ lam, int stride, int off,
                                                     double* __restrict__ part) {
  __shared__ double sh[kT / 64];
  const uint32_t tile = tiles[blockIdx.x];
  double l = 0;
  if (MODE == 2) l = lam[size_t(tile_sector[tile]) * stride + off];
  double s = 0;
#pragma unroll
  for (int i = 0; i < kRPT; ++i) {
    const size_t r = row_of(tile, i);
    if (MODE == 0) s += nsq(x[r]);
    if (MODE == 1) s += cmulc(x[r], y[r]).x;
    if (MODE == 2) {
      const double2 a = x[r], b = y[r];
      s += nsq(make_double2(b.x - l * a.x, b.y - l * a.y));
    }
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[tile] = s;
}

// out[sector · stride + off] = Σ part over the sector's tiles (its square
// root when root), one wave per listed sector
__global__ __launch_bounds__(kT) void kb_seg_sum(const uint32_t* __restrict__ secs, int nsecs,
                                                 const Seg* __restrict__ seg, const double* __restrict__ part,
                                                 double* out, int stride, int off, int root) {
  const int item = blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (item >= nsecs) return;
  const uint32_t s = secs[item];
  const Seg g = seg[s];
  double t = 0;
  for (uint32_t b = lane; b < g.ntiles; b += 64) t += part[g.tile0 + b];
  t = wave_sum(t);
  if (lane == 0) out[size_t(s) * stride + off] = root ? sqrt(t) : t;
}

// x ← x / sqrt(nrm[sector · stride + off]) on the listed tiles
__global__ __launch_bounds__(kT) void kb_scale(const uint32_t* __restrict__ tiles,
                                               const uint32_t* __restrict__ tile_sector,
                                               const double* __restrict__ nrm, int stride, int off,
                                               double2* __restrict__ x) {
  const uint32_t tile = tiles[blockIdx.x];
  const double inv = 1.0 / sqrt(nrm[size_t(tile_sector[tile]) * stride + off]);
#pragma unroll
  for (int i = 0; i < kRPT; ++i) {
    const size_t r = row_of(tile, i);
    const double2 a = x[r];
    x[r] = make_double2(a.x * inv, a.y * inv);
  }
}

// ------------------------------------------------------- Lanczos phases ---

// The three-term part of the first pass (src/solve/mod.rs:1040-1048):
// w ← w − α_j v_j − β_{j−1} v_{j−1}, with the per-tile partials of ‖w₀‖².
__global__ __launch_bounds__(kT) void kb_three(const uint32_t* __restrict__ tiles,
                                               const uint32_t* __restrict__ tile_sector,
                                               const double2* __restrict__ vj, const double2* __restrict__ vjm1,
                                               const double* __restrict__ alpha, const double* __restrict__ beta,
                                               int m, int j, double2* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[kT / 64];
  const uint32_t tile = tiles[blockIdx.x];
  const size_t s = tile_sector[tile];
  const double a = alpha[s * m + j];
  const double b = j > 0 ? beta[s * m + j - 1] : 0.0;
  double acc = 0;
#pragma unroll
  for (int i = 0; i < kRPT; ++i) {
    const size_t r = row_of(tile, i);
    double2 x = w[r];
    const double2 v = vj[r];
    x.x -= a * v.x;
    x.y -= a * v.y;
    if (j > 0) {
      const double2 u = vjm1[r];
      x.x -= b * u.x;
      x.y -= b * u.y;
    }
    w[r] = x;
    acc += nsq(x);
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[tile] = acc;
}

// Overlaps of one tile: partial[tile · ld + c] = Σ_rows conj(V_c) w for
// c < ncols. Columns go kCT at a time; each wave's sums collect in LDS
// (sh: 4 waves × ncols) and are added in wave order at the end.
__device__ __forceinline__ void dots_tile(uint32_t tile, const double2* __restrict__ V, size_t T, int ncols,
                                          const double2* __restrict__ w, double2* __restrict__ partial, int ld,
                                          double2* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double2 wv[kRPT];
  size_t rr[kRPT];
#pragma unroll
  for (int i = 0; i < kRPT; ++i) {
    rr[i] = row_of(tile, i);
    wv[i] = w[rr[i]];
  }
  for (int c0 = 0; c0 < ncols; c0 += kCT) {
    double2 acc[kCT];
#pragma unroll
    for (int c = 0; c < kCT; ++c) {
      acc[c] = make_double2(0.0, 0.0);
      const size_t base = size_t(min(c0 + c, ncols - 1)) * T;
#pragma unroll
      for (int i = 0; i < kRPT; ++i) acc[c] = cadd(acc[c], cmulc(V[base + rr[i]], wv[i]));
    }
#pragma unroll
    for (int c = 0; c < kCT; ++c) {
      const double2 t = wave_sum(acc[c]);
      if (lane == 0 && c0 + c < ncols) sh[wid * ncols + c0 + c] = t;
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < ncols; c += kT) {
    double2 t = sh[c];
    for (int q = 1; q < kT / 64; ++q) t = cadd(t, sh[q * ncols + c]);
    partial[size_t(tile) * ld + c] = t;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kT) void kb_dots(const uint32_t* __restrict__ tiles, const double2* __restrict__ V,
                                              size_t T, int ncols, const double2* __restrict__ w,
                                              double2* __restrict__ partial, int ld) {
  extern __shared__ double2 sh_dots[];
  dots_tile(tiles[blockIdx.x], V, T, ncols, w, partial, ld, sh_dots);
}

// w ← w − V h_s over one tile (h_s: the sector's ncols overlaps), with the
// per-tile partial of ‖w‖²
__device__ __forceinline__ void sub_tile(uint32_t tile, const double2* __restrict__ V, size_t T, int ncols,
                                         const double2* __restrict__ hs, double2* __restrict__ w,
                                         double* __restrict__ part, double2* shh, double* shr) {
  for (int c = threadIdx.x; c < ncols; c += kT) shh[c] = hs[c];
  __syncthreads();
  double2 acc[kRPT];
  size_t rr[kRPT];
#pragma unroll
  for (int i = 0; i < kRPT; ++i) {
    rr[i] = row_of(tile, i);
    acc[i] = make_double2(0.0, 0.0);
  }
#pragma unroll 2
  for (int c = 0; c < ncols; ++c) {
    const double2 hc = shh[c];
    const double2* col = V + size_t(c) * T;
#pragma unroll
    for (int i = 0; i < kRPT; ++i) acc[i] = cadd(acc[i], cmul(col[rr[i]], hc));
  }
  double s = 0;
#pragma unroll
  for (int i = 0; i < kRPT; ++i) {
    const double2 x = w[rr[i]];
    const double2 v = make_double2(x.x - acc[i].x, x.y - acc[i].y);
    w[rr[i]] = v;
    s += nsq(v);
  }
  s = block_sum(s, shr);  // ends with a barrier: shh may be reloaded
  if (threadIdx.x == 0) part[tile] = s;
}

__global__ __launch_bounds__(kT) void kb_sub(const uint32_t* __restrict__ tiles,
                                             const uint32_t* __restrict__ tile_sector, const double2* __restrict__ V,
                                             size_t T, int ncols, const double2* __restrict__ h, int ld,
                                             double2* __restrict__ w, double* __restrict__ part) {
  extern __shared__ double2 sh_sub[];
  __shared__ double shr[kT / 64];
  const uint32_t tile = tiles[blockIdx.x];
  sub_tile(tile, V, T, ncols, h + size_t(tile_sector[tile]) * ld, w, part, sh_sub, shr);
}

// Per-sector sums of the overlap partials, one wave per (sector, column):
// h[s · ld + c] for c < ncols; with pn, column ncols is ‖w₀‖² = Σ pn into
// w0n[s]. Sectors come from `secs` (count nsecs) or, when list != nullptr,
// from the second-pass list (count list[0], sectors list[1..]).
__global__ __launch_bounds__(kT) void kb_hred(const uint32_t* __restrict__ secs, int nsecs,
                                              const uint32_t* __restrict__ list, const Seg* __restrict__ seg,
                                              const double2* __restrict__ partial, int ld, int ncols,
                                              double2* __restrict__ h, const double* __restrict__ pn,
                                              double* __restrict__ w0n) {
  const int lane = threadIdx.x & 63;
  if (list) {
    nsecs = int(list[0]);
    secs = list + 1;
  }
  const int per = ncols + (pn ? 1 : 0);
  const int total = nsecs * per;
  for (int item = blockIdx.x * (kT / 64) + (threadIdx.x >> 6); item < total; item += gridDim.x * (kT / 64)) {
    const uint32_t s = secs[item / per];
    const int c = item % per;
    const Seg g = seg[s];
    if (c < ncols) {
      double2 t = make_double2(0.0, 0.0);
      for (uint32_t b = lane; b < g.ntiles; b += 64) t = cadd(t, partial[size_t(g.tile0 + b) * ld + c]);
      t = wave_sum(t);
      if (lane == 0) h[size_t(s) * ld + c] = t;
    } else {
      double t = 0;
      for (uint32_t b = lane; b < g.ntiles; b += 64) t += pn[g.tile0 + b];
      t = wave_sum(t);
      if (lane == 0) w0n[s] = t;
    }
  }
}

// DGKS decision of step j for every running sector (one block): ‖w‖² = Σ pn;
// a second pass iff ‖w₀‖ > eps and ‖w‖ <= η‖w₀‖. Otherwise β_j = ‖w‖. The
// sectors needing a second pass are listed in running order:
// list[0] = count, list[1 + i] = sector, list[1 + kMaxSectors + i] = first
// work tile of that sector (prefix of tile counts; entry count = total).
__global__ __launch_bounds__(kT) void kb_decide(const uint32_t* __restrict__ secs, int nsecs,
                                                const Seg* __restrict__ seg, const double* __restrict__ pn,
                                                const double* __restrict__ w0n, int m, int j, double eta,
                                                double* __restrict__ beta, int* __restrict__ hist,
                                                uint32_t* __restrict__ list) {
  __shared__ uint8_t flag[kMaxSectors];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int i = wid; i < nsecs; i += kT / 64) {
    const uint32_t s = secs[i];
    const Seg g = seg[s];
    double t = 0;
    for (uint32_t b = lane; b < g.ntiles; b += 64) t += pn[g.tile0 + b];
    t = wave_sum(t);
    if (lane == 0) {
      const double na = sqrt(t), nb = sqrt(w0n[s]);
      const bool second = nb > kEps && na <= eta * nb;
      flag[i] = second ? 1 : 0;
      hist[size_t(s) * m + j] = second ? 1 : 0;
      if (!second) beta[size_t(s) * m + j] = na;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t cnt = 0, pref = 0;
    for (int i = 0; i < nsecs; ++i) {
      if (!flag[i]) continue;
      const uint32_t s = secs[i];
      list[1 + cnt] = s;
      list[1 + kMaxSectors + cnt] = pref;
      pref += seg[s].ntiles;
      ++cnt;
    }
    list[1 + kMaxSectors + cnt] = pref;
    list[0] = cnt;
  }
}

// work item -> (sector, tile) of the second-pass list
__device__ __forceinline__ void listed_tile(const uint32_t* list, const Seg* seg, uint32_t item, uint32_t& s,
                                            uint32_t& tile) {
  const uint32_t cnt = list[0];
  const uint32_t* pref = list + 1 + kMaxSectors;
  uint32_t lo = 0, hi = cnt;  // pref[lo] <= item < pref[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (pref[mid] <= item) lo = mid;
    else hi = mid;
  }
  s = list[1 + lo];
  tile = seg[s].tile0 + (item - pref[lo]);
}

__global__ __launch_bounds__(kT) void kb_dots2(const uint32_t* __restrict__ list, const Seg* __restrict__ seg,
                                               const double2* __restrict__ V, size_t T, int ncols,
                                               const double2* __restrict__ w, double2* __restrict__ partial, int ld) {
  extern __shared__ double2 sh_dots2[];
  const uint32_t total = list[1 + kMaxSectors + list[0]];
  for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
    uint32_t s, tile;
    listed_tile(list, seg, item, s, tile);
    dots_tile(tile, V, T, ncols, w, partial, ld, sh_dots2);
  }
}

__global__ __launch_bounds__(kT) void kb_sub2(const uint32_t* __restrict__ list, const Seg* __restrict__ seg,
                                              const double2* __restrict__ V, size_t T, int ncols,
                                              const double2* __restrict__ h, int ld, double2* __restrict__ w,
                                              double* __restrict__ part) {
  extern __shared__ double2 sh_sub2[];
  __shared__ double shr[kT / 64];
  const uint32_t total = list[1 + kMaxSectors + list[0]];
  for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
    uint32_t s, tile;
    listed_tile(list, seg, item, s, tile);
    sub_tile(tile, V, T, ncols, h + size_t(s) * ld, w, part, sh_sub2, shr);
  }
}

// after a second pass: β_j = ‖w‖ for the listed sectors
__global__ __launch_bounds__(kT) void kb_decide2(const uint32_t* __restrict__ list, const Seg* __restrict__ seg,
                                                 const double* __restrict__ pn, int m, int j,
                                                 double* __restrict__ beta) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int cnt = int(list[0]);
  for (int i = wid; i < cnt; i += kT / 64) {
    const uint32_t s = list[1 + i];
    const Seg g = seg[s];
    double t = 0;
    for (uint32_t b = lane; b < g.ntiles; b += 64) t += pn[g.tile0 + b];
    t = wave_sum(t);
    if (lane == 0) beta[size_t(s) * m + j] = sqrt(t);
  }
}

// v_{j+1} = w / β_j
__global__ __launch_bounds__(kT) void kb_next(const uint32_t* __restrict__ tiles,
                                              const uint32_t* __restrict__ tile_sector, const double2* __restrict__ w,
                                              const double* __restrict__ beta, int m, int j,
                                              double2* __restrict__ out) {
  const uint32_t tile = tiles[blockIdx.x];
  const double inv = 1.0 / beta[size_t(tile_sector[tile]) * m + j];
#pragma unroll
  for (int i = 0; i < kRPT; ++i) {
    const size_t r = row_of(tile, i);
    const double2 x = w[r];
    out[r] = make_double2(x.x * inv, x.y * inv);
  }
}

// Ritz vectors Y_q = Σ_c V_c C_s[q][c] for q in [q0, q0 + K), C_s real
// (tridiagonal eigenvectors) at coef + coef_off[s], layout [q][c] with the
// sector's window size sizes[s]
template <int K>
__global__ __launch_bounds__(kT) void kb_ritz(const uint32_t* __restrict__ tiles,
                                              const uint32_t* __restrict__ tile_sector, const double2* __restrict__ V,
                                              size_t T, const double* __restrict__ coef,
                                              const uint32_t* __restrict__ coef_off, const int* __restrict__ sizes,
                                              int q0, double2* __restrict__ Y) {
  extern __shared__ double shc[];
  const uint32_t tile = tiles[blockIdx.x];
  const uint32_t s = tile_sector[tile];
  const int sz = sizes[s];
  const double* cs = coef + coef_off[s];
  for (int i = threadIdx.x; i < K * sz; i += kT) shc[i] = cs[size_t(q0) * sz + i];
  __syncthreads();
  constexpr int H = kRPT / 2;
  for (int half = 0; half < 2; ++half) {
    double2 acc[H][K];
    size_t rr[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
      rr[i] = row_of(tile, half * H + i);
#pragma unroll
      for (int q = 0; q < K; ++q) acc[i][q] = make_double2(0.0, 0.0);
    }
    for (int c = 0; c < sz; ++c) {
      const double2* col = V + size_t(c) * T;
#pragma unroll
      for (int i = 0; i < H; ++i) {
        const double2 v = col[rr[i]];
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const double cq = shc[q * sz + c];
          acc[i][q].x += v.x * cq;
          acc[i][q].y += v.y * cq;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < H; ++i)
#pragma unroll
      for (int q = 0; q < K; ++q) Y[size_t(q0 + q) * T + rr[i]] = acc[i][q];
  }
}

int groups(int items, int per_block) { return (items + per_block - 1) / per_block; }

// f(i) for i < count on up to 16 host threads (the per-sector tridiagonal
// work of a read-back; each i touches only its own sector's state)
template <typename F>
void parallel_for(int count, F f) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int nth = std::min(count, int(std::max(1u, std::min(16u, hw ? hw : 1u))));
  if (nth <= 1) {
    for (int i = 0; i < count; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  std::vector<std::exception_ptr> err(static_cast<size_t>(nth));
  auto work = [&](int tid) {
    try {
      for (int i = tid; i < count; i += nth) f(i);
    } catch (...) {
      err[size_t(tid)] = std::current_exception();
    }
  };
  for (int t = 1; t < nth; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

// Host side of one batch.
class Batch {
 public:
  Batch(const std::vector<Operator*>& ops, const EigshOptions& o, const EigshControls& ctl, int m, hipStream_t s)
      : ops_(ops), o_(o), ctl_(ctl), m_(m), ld_(m + 1), k_(o.eigenpairs), s_(s) {
    S_ = int(ops.size());
    std::vector<size_t> offsets;
    uint32_t tile = 0;
    for (int i = 0; i < S_; ++i) {
      const size_t n = ops[i]->dim();
      const uint32_t nt = uint32_t((n + kTile - 1) / kTile);
      seg_.push_back({tile, nt, uint32_t(n), 0});
      offsets.push_back(size_t(tile) * kTile);
      for (uint32_t t = 0; t < nt; ++t) tile_sector_.push_back(uint32_t(i));
      tile += nt;
    }
    tiles_ = tile;
    T_ = size_t(tiles_) * kTile;
    action_ = std::make_unique<SectorBatchAction>(ops, offsets, s_);
    // memory: Krylov basis, w, Ritz block, H·y, partials
    const size_t need = (size_t(m_) + 2 + size_t(k_)) * T_ * 16 + size_t(tiles_) * ld_ * 16;
    size_t free_b = 0, total_b = 0;
    QMBED_HIP(hipMemGetInfo(&free_b, &total_b));
    if (need > free_b / 10 * 9)
      throw unsupported_backend("sector batch: the Krylov window of all sectors does not fit in device memory");
    V_.resize(2 * T_ * size_t(m_));
    w_.resize(2 * T_);
    Y_.resize(2 * T_ * size_t(k_));
    Z_.resize(2 * T_);
    // Every column of V and Y is written over whole tiles of the sectors
    // that later read it; w and Z receive H·v output on live rows only, so
    // their padding rows must start (and stay) zero.
    QMBED_HIP(hipMemsetAsync(w_.data(), 0, w_.size() * 8, s_));
    QMBED_HIP(hipMemsetAsync(Z_.data(), 0, Z_.size() * 8, s_));
    partial_.resize(2 * size_t(tiles_) * ld_);
    pa_.resize(tiles_);
    pn_.resize(tiles_);
    coef_.resize(2 * size_t(S_) * m_);  // α then β
    hist_.resize(size_t(S_) * m_);
    h_.resize(2 * size_t(S_) * ld_);
    w0n_.resize(S_);
    slot_.resize(3 * size_t(S_) * k_ + 64);
    list_.resize(2 + 2 * size_t(kMaxSectors));
    seg_d_.resize(std::max<size_t>(1, seg_.size() * sizeof(Seg)));
    QMBED_HIP(hipMemcpy(seg_d_.data(), seg_.data(), seg_.size() * sizeof(Seg), hipMemcpyHostToDevice));
    ts_d_.resize(std::max<size_t>(1, tile_sector_.size()));
    QMBED_HIP(hipMemcpy(ts_d_.data(), tile_sector_.data(), tile_sector_.size() * 4, hipMemcpyHostToDevice));
    uint32_t hv_total = 0;
    for (int i = 0; i < S_; ++i) hv_total += action_->blocks_of(i);
    act_tiles_.resize(tiles_);
    act_secs_.resize(S_);
    act_hv_.resize(hv_total);
    r_tiles_.resize(tiles_);
    r_secs_.resize(S_);
    r_hv_.resize(hv_total);
    coef_off_.resize(S_);
    sizes_.resize(S_);
    ritz_coef_.resize(size_t(S_) * k_ * m_ + 1);
    // pinned staging: lists (3 regions of tiles + S + hv words), read-back
    stage_words_ = 2 * (size_t(tiles_) + S_ + hv_total) + 2 * size_t(S_) + 16;
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage_), stage_words_ * 4, hipHostMallocDefault));
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&ritz_stage_), (size_t(S_) * k_ * m_ + 1) * 8,
                            hipHostMallocDefault));
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&back_), (2 * size_t(S_) * m_) * 8, hipHostMallocDefault));
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&back_hist_), size_t(S_) * m_ * 4, hipHostMallocDefault));
    QMBED_HIP(hipHostMalloc(reinterpret_cast<void**>(&slot_host_), slot_.size() * 8, hipHostMallocDefault));
    QMBED_HIP(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
    state_.resize(S_);
    for (int i = 0; i < S_; ++i) state_[i].next_check = std::max(k_ + 1, ctl_.check_every);
  }
  ~Batch() {
    (void)hipStreamSynchronize(s_);
    if (stage_) (void)hipHostFree(stage_);
    if (ritz_stage_) (void)hipHostFree(ritz_stage_);
    if (back_) (void)hipHostFree(back_);
    if (back_hist_) (void)hipHostFree(back_hist_);
    if (slot_host_) (void)hipHostFree(slot_host_);
    if (ev_) (void)hipEventDestroy(ev_);
  }
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  struct SectorState {
    bool running = true;
    bool accepted = false;
    bool fallback = false;
    int size = 0;
    int next_check = 0;
    long matvecs = 0;  // H·v launched for the sector (steps while running + Rayleigh)
    long steps_seen = 0;
    std::vector<double> alpha, beta;
    double last_beta = 0;
    bool broke = false;
    Eigensystem es;
    std::vector<std::vector<cplx>> vectors;
  };

  void run(bool want_vectors) {
    want_vectors_ = want_vectors;
    std::vector<int> all(S_);
    std::iota(all.begin(), all.end(), 0);
    upload_running(all);
    start();
    constexpr int kLag = 2;
    const int every = std::max(1, ctl_.check_every);
    int launched = 0, posted = -1;
    bool pending = false;
    while (!running_.empty()) {
      if (launched < m_) {
        step(launched);
        ++launched;
        ++steps_;
        for (int i : running_) ++state_[i].matvecs;
      }
      if (!pending && (launched % every == 0 || launched == m_)) {
        post();
        posted = launched;
        pending = true;
      }
      if (!pending || (launched < posted + kLag && launched < m_)) continue;
      pending = false;
      collect(posted);
    }
  }

  const SectorState& state(int i) const { return state_[size_t(i)]; }
  long steps() const { return steps_; }
  int ritz_rounds() const { return ritz_rounds_; }

 private:
  double2* col(int j) const { return reinterpret_cast<double2*>(V_.data()) + T_ * size_t(j); }
  double2* w() const { return reinterpret_cast<double2*>(w_.data()); }
  double* alpha() const { return coef_.data(); }
  double* beta() const { return coef_.data() + size_t(S_) * m_; }
  const Seg* seg() const { return reinterpret_cast<const Seg*>(seg_d_.data()); }

  // list regions of the pinned staging area: running lists, then Ritz lists
  void upload_list(const std::vector<uint32_t>& v, DeviceArray<uint32_t>& dst, size_t& cursor) {
    if (v.empty()) return;
    if (cursor + v.size() > stage_words_) throw internal_state("sector batch staging area exceeded");
    std::copy(v.begin(), v.end(), stage_ + cursor);
    QMBED_HIP(hipMemcpyAsync(dst.data(), stage_ + cursor, v.size() * 4, hipMemcpyHostToDevice, s_));
    cursor += v.size();
  }

  void lists_of(const std::vector<int>& secs, std::vector<uint32_t>& tiles, std::vector<uint32_t>& ids,
                std::vector<uint32_t>& hv) const {
    tiles.clear();
    ids.clear();
    hv.clear();
    for (int i : secs) {
      ids.push_back(uint32_t(i));
      for (uint32_t t = 0; t < seg_[i].ntiles; ++t) tiles.push_back(seg_[i].tile0 + t);
      for (uint32_t b = 0; b < action_->blocks_of(i); ++b) hv.push_back(uint32_t(i) << 16 | b);
    }
  }

  // The running lists are rewritten only after a stream synchronization (the
  // start, or a Ritz round), so the staging words are no longer being read.
  void upload_running(const std::vector<int>& secs) {
    running_ = secs;
    std::vector<uint32_t> tiles, ids, hv;
    lists_of(secs, tiles, ids, hv);
    size_t cursor = 0;
    upload_list(tiles, act_tiles_, cursor);
    upload_list(ids, act_secs_, cursor);
    upload_list(hv, act_hv_, cursor);
    n_tiles_ = int(tiles.size());
    n_secs_ = int(ids.size());
    n_hv_ = int(hv.size());
  }

  void start() {
    const uint32_t* ts = ts_d_.data();
    hipLaunchKernelGGL(kb_start, dim3(n_tiles_), dim3(kT), 0, s_, act_tiles_.data(), ts, seg(), o_.seed, col(0));
    hipLaunchKernelGGL(kb_tile_reduce<0>, dim3(n_tiles_), dim3(kT), 0, s_, act_tiles_.data(), ts, col(0), nullptr,
                       nullptr, 0, 0, pn_.data());
    hipLaunchKernelGGL(kb_seg_sum, dim3(groups(n_secs_, kT / 64)), dim3(kT), 0, s_, act_secs_.data(), n_secs_, seg(),
                       pn_.data(), w0n_.data(), 1, 0, 0);
    hipLaunchKernelGGL(kb_scale, dim3(n_tiles_), dim3(kT), 0, s_, act_tiles_.data(), ts, w0n_.data(), 1, 0, col(0));
    QMBED_HIP(hipGetLastError());
  }

  void step(int j) {
    const uint32_t* ts = ts_d_.data();
    const uint32_t* tl = act_tiles_.data();
    const uint32_t* sl = act_secs_.data();
    const int ncols = j + 1;
    double2* part = reinterpret_cast<double2*>(partial_.data());
    double2* h = reinterpret_cast<double2*>(h_.data());
    action_->apply(act_hv_.data(), n_hv_, reinterpret_cast<const double*>(col(j)), w_.data(), s_);
    hipLaunchKernelGGL(kb_tile_reduce<1>, dim3(n_tiles_), dim3(kT), 0, s_, tl, ts, col(j), w(), nullptr, 0, 0,
                       pa_.data());
    hipLaunchKernelGGL(kb_seg_sum, dim3(groups(n_secs_, kT / 64)), dim3(kT), 0, s_, sl, n_secs_, seg(), pa_.data(),
                       alpha(), m_, j, 0);
    hipLaunchKernelGGL(kb_three, dim3(n_tiles_), dim3(kT), 0, s_, tl, ts, col(j), j > 0 ? col(j - 1) : nullptr,
                       alpha(), beta(), m_, j, w(), pn_.data());
    const size_t sm_dots = size_t(kT / 64) * ncols * sizeof(double2);
    hipLaunchKernelGGL(kb_dots, dim3(n_tiles_), dim3(kT), sm_dots, s_, tl, col(0), T_, ncols, w(), part, ld_);
    hipLaunchKernelGGL(kb_hred, dim3(groups(n_secs_ * (ncols + 1), kT / 64)), dim3(kT), 0, s_, sl, n_secs_, nullptr,
                       seg(), part, ld_, ncols, h, pn_.data(), w0n_.data());
    const size_t sm_sub = size_t(ncols) * sizeof(double2);
    hipLaunchKernelGGL(kb_sub, dim3(n_tiles_), dim3(kT), sm_sub, s_, tl, ts, col(0), T_, ncols, h, ld_, w(),
                       pn_.data());
    hipLaunchKernelGGL(kb_decide, dim3(1), dim3(kT), 0, s_, sl, n_secs_, seg(), pn_.data(), w0n_.data(), m_, j,
                       ctl_.dgks_eta, beta(), hist_.data(), list_.data());
    // second pass over the listed sectors (no-ops when the list is empty)
    hipLaunchKernelGGL(kb_dots2, dim3(kSecondGrid), dim3(kT), sm_dots, s_, list_.data(), seg(), col(0), T_, ncols,
                       w(), part, ld_);
    hipLaunchKernelGGL(kb_hred, dim3(kSecondGrid / 4), dim3(kT), 0, s_, nullptr, 0, list_.data(), seg(), part, ld_,
                       ncols, h, nullptr, nullptr);
    hipLaunchKernelGGL(kb_sub2, dim3(kSecondGrid), dim3(kT), sm_sub, s_, list_.data(), seg(), col(0), T_, ncols, h,
                       ld_, w(), pn_.data());
    hipLaunchKernelGGL(kb_decide2, dim3(1), dim3(kT), 0, s_, list_.data(), seg(), pn_.data(), m_, j, beta());
    if (j + 1 < m_)
      hipLaunchKernelGGL(kb_next, dim3(n_tiles_), dim3(kT), 0, s_, tl, ts, w(), beta(), m_, j, col(j + 1));
    QMBED_HIP(hipGetLastError());
  }

  void post() {
    QMBED_HIP(hipMemcpyAsync(back_, coef_.data(), coef_.size() * 8, hipMemcpyDeviceToHost, s_));
    QMBED_HIP(hipMemcpyAsync(back_hist_, hist_.data(), hist_.size() * 4, hipMemcpyDeviceToHost, s_));
    QMBED_HIP(hipEventRecord(ev_, s_));
  }

  // Parse the read-back of the first L steps for every running sector and
  // decide which sectors take the Ritz test now.
  void collect(int L) {
    QMBED_HIP(hipEventSynchronize(ev_));
    std::vector<int> test, check;
    std::vector<char> ended(S_, 0);
    for (int i : running_) {
      SectorState& st = state_[i];
      const double* a = back_ + size_t(i) * m_;
      const double* b = back_ + size_t(S_) * m_ + size_t(i) * m_;
      st.alpha.clear();
      st.beta.clear();
      st.broke = false;
      bool bad = false;
      int size = L;
      for (int j = 0; j < L; ++j) {
        if (!std::isfinite(a[j]) || !std::isfinite(b[j])) {
          bad = true;
          break;
        }
        if (b[j] <= 1.0e-14) {
          size = j + 1;
          st.broke = true;
          break;
        }
      }
      if (bad) {  // left to the single-sector solver, which reports it
        st.running = false;
        st.fallback = true;
        continue;
      }
      for (int j = 0; j < size; ++j) {
        st.alpha.push_back(a[j]);
        if (j + 1 < size) st.beta.push_back(b[j]);
      }
      st.last_beta = b[size - 1];
      st.size = size;
      const bool end = st.broke || size == m_;
      ended[i] = end ? 1 : 0;
      if (end) {
        test.push_back(i);
        continue;
      }
      if (!ctl_.early_stop || size < st.next_check || size <= k_) continue;
      check.push_back(i);
    }
    // residual estimates of the sectors due for a check, in parallel
    std::vector<char> conv(check.size(), 0);
    parallel_for(int(check.size()), [&](int q) { conv[size_t(q)] = estimates_converged(state_[check[q]]) ? 1 : 0; });
    for (size_t q = 0; q < check.size(); ++q) {
      if (conv[q]) test.push_back(check[q]);
      else state_[check[q]].next_check = state_[check[q]].size + ctl_.check_every;
    }
    std::sort(test.begin(), test.end());
    std::vector<int> keep;
    if (!test.empty()) ritz(test);
    for (int i : running_) {
      SectorState& st = state_[i];
      if (!st.running) continue;
      if (std::find(test.begin(), test.end(), i) != test.end()) {
        if (st.accepted) st.running = false;
        else if (ended[i]) {
          st.running = false;
          st.fallback = true;
        } else {
          st.next_check = st.size + ctl_.check_every;
        }
      }
      if (st.running) keep.push_back(i);
    }
    if (keep.size() != running_.size()) {
      QMBED_HIP(hipStreamSynchronize(s_));
      upload_running(keep);
    }
  }

  bool estimates_converged(const SectorState& st) const {
    const int sz = st.size;
    const auto vals = selected_ritz_values(st.alpha, st.beta, o_.target, k_, o_.shift);
    if (int(vals.size()) < k_) return false;
    for (double val : vals) {
      auto v = tridiagonal_eigvec(st.alpha, st.beta, val);
      const double est = st.last_beta * std::fabs(v[sz - 1]);
      if (!std::isfinite(est) || est > 0.5 * o_.tolerance * std::max(std::fabs(val), 1.0)) return false;
    }
    return true;
  }

  // Ritz pairs of the listed sectors, FP64 Rayleigh quotients and residual
  // norms (one synchronization), acceptance as in the single-sector window.
  void ritz(const std::vector<int>& secs) {
    ++ritz_rounds_;
    std::vector<uint32_t> offs(S_, 0);
    std::vector<int> sizes(S_, 0);
    std::vector<std::vector<double>> vals(S_);
    std::vector<std::vector<std::vector<double>>> vecs(S_);
    // selected Ritz pairs of the tridiagonal matrices, in parallel
    parallel_for(int(secs.size()), [&](int q) {
      const int i = secs[size_t(q)];
      vals[i] = selected_ritz_values(state_[i].alpha, state_[i].beta, o_.target, k_, o_.shift, &vecs[i]);
    });
    size_t cursor = 0;
    for (int i : secs) {
      const int sz = state_[i].size;
      offs[i] = uint32_t(cursor);
      sizes[i] = sz;
      for (int q = 0; q < k_; ++q) {
        const bool have = q < int(vecs[i].size());
        for (int c = 0; c < sz; ++c) ritz_stage_[cursor + size_t(q) * sz + c] = have ? vecs[i][q][size_t(c)] : 0.0;
      }
      cursor += size_t(k_) * sz;
    }
    QMBED_HIP(hipMemcpyAsync(ritz_coef_.data(), ritz_stage_, cursor * 8, hipMemcpyHostToDevice, s_));
    std::vector<uint32_t> tiles, ids, hv;
    lists_of(secs, tiles, ids, hv);
    // Ritz lists follow the running lists in the staging area
    size_t lc = size_t(tiles_) + S_ + act_hv_.size();
    for (int i = 0; i < S_; ++i) stage_[lc + i] = offs[i];
    QMBED_HIP(hipMemcpyAsync(coef_off_.data(), stage_ + lc, size_t(S_) * 4, hipMemcpyHostToDevice, s_));
    lc += S_;
    for (int i = 0; i < S_; ++i) reinterpret_cast<int*>(stage_)[lc + i] = sizes[i];
    QMBED_HIP(hipMemcpyAsync(sizes_.data(), stage_ + lc, size_t(S_) * 4, hipMemcpyHostToDevice, s_));
    lc += S_;
    upload_list(tiles, r_tiles_, lc);
    upload_list(ids, r_secs_, lc);
    upload_list(hv, r_hv_, lc);
    const int nt = int(tiles.size()), ns = int(ids.size()), nh = int(hv.size());
    const uint32_t* ts = ts_d_.data();
    int maxsz = 0;
    for (int i : secs) maxsz = std::max(maxsz, sizes[i]);
    double2* Y = reinterpret_cast<double2*>(Y_.data());
    for (int q0 = 0; q0 < k_; q0 += 4) {
      const int kk = std::min(4, k_ - q0);
      const size_t sm = size_t(kk) * maxsz * 8;
      switch (kk) {
        case 1: hipLaunchKernelGGL(kb_ritz<1>, dim3(nt), dim3(kT), sm, s_, r_tiles_.data(), ts, col(0), T_,
                                   ritz_coef_.data(), coef_off_.data(), sizes_.data(), q0, Y); break;
        case 2: hipLaunchKernelGGL(kb_ritz<2>, dim3(nt), dim3(kT), sm, s_, r_tiles_.data(), ts, col(0), T_,
                                   ritz_coef_.data(), coef_off_.data(), sizes_.data(), q0, Y); break;
        case 3: hipLaunchKernelGGL(kb_ritz<3>, dim3(nt), dim3(kT), sm, s_, r_tiles_.data(), ts, col(0), T_,
                                   ritz_coef_.data(), coef_off_.data(), sizes_.data(), q0, Y); break;
        default: hipLaunchKernelGGL(kb_ritz<4>, dim3(nt), dim3(kT), sm, s_, r_tiles_.data(), ts, col(0), T_,
                                    ritz_coef_.data(), coef_off_.data(), sizes_.data(), q0, Y); break;
      }
    }
    // slots: [0, Sk) ‖y‖², [Sk, 2Sk) λ, [2Sk, 3Sk) residual norm
    double* nrm = slot_.data();
    double* lam = nrm + size_t(S_) * k_;
    double* res = lam + size_t(S_) * k_;
    double2* Z = reinterpret_cast<double2*>(Z_.data());
    const dim3 sg(groups(ns, kT / 64));
    for (int q = 0; q < k_; ++q) {
      double2* y = Y + T_ * size_t(q);
      hipLaunchKernelGGL(kb_tile_reduce<0>, dim3(nt), dim3(kT), 0, s_, r_tiles_.data(), ts, y, nullptr, nullptr, 0, 0,
                         pn_.data());
      hipLaunchKernelGGL(kb_seg_sum, sg, dim3(kT), 0, s_, r_secs_.data(), ns, seg(), pn_.data(), nrm, k_, q, 0);
      hipLaunchKernelGGL(kb_scale, dim3(nt), dim3(kT), 0, s_, r_tiles_.data(), ts, nrm, k_, q, y);
      action_->apply(r_hv_.data(), nh, reinterpret_cast<const double*>(y), Z_.data(), s_);
      hipLaunchKernelGGL(kb_tile_reduce<1>, dim3(nt), dim3(kT), 0, s_, r_tiles_.data(), ts, y, Z, nullptr, 0, 0,
                         pa_.data());
      hipLaunchKernelGGL(kb_seg_sum, sg, dim3(kT), 0, s_, r_secs_.data(), ns, seg(), pa_.data(), lam, k_, q, 0);
      hipLaunchKernelGGL(kb_tile_reduce<2>, dim3(nt), dim3(kT), 0, s_, r_tiles_.data(), ts, y, Z, lam, k_, q,
                         pn_.data());
      hipLaunchKernelGGL(kb_seg_sum, sg, dim3(kT), 0, s_, r_secs_.data(), ns, seg(), pn_.data(), res, k_, q, 1);
    }
    QMBED_HIP(hipGetLastError());
    QMBED_HIP(hipMemcpyAsync(slot_host_, slot_.data(), 3 * size_t(S_) * k_ * 8, hipMemcpyDeviceToHost, s_));
    QMBED_HIP(hipStreamSynchronize(s_));
    for (int i : secs) {
      SectorState& st = state_[i];
      st.matvecs += k_;
      struct Cand {
        double value, residual;
        int column;
      };
      std::vector<Cand> cands;
      for (int q = 0; q < int(vals[i].size()); ++q)
        cands.push_back({slot_host_[size_t(S_) * k_ + size_t(i) * k_ + q],
                         slot_host_[2 * size_t(S_) * k_ + size_t(i) * k_ + q], q});
      const SpectrumTarget t = o_.target == SpectrumTarget::BothEnds ? SpectrumTarget::SmallestAlgebraic : o_.target;
      std::stable_sort(cands.begin(), cands.end(), [&](const Cand& a, const Cand& b) {
        switch (t) {
          case SpectrumTarget::LargestAlgebraic: return b.value < a.value;
          case SpectrumTarget::SmallestMagnitude: return std::fabs(a.value) < std::fabs(b.value);
          case SpectrumTarget::LargestMagnitude: return std::fabs(b.value) < std::fabs(a.value);
          default: return a.value < b.value;
        }
      });
      bool ok = int(cands.size()) == k_;
      for (auto& c : cands)
        if (!(c.residual <= o_.tolerance * std::max(std::fabs(c.value), 1.0))) ok = false;
      if (!ok) continue;
      st.accepted = true;
      Eigensystem& es = st.es;
      for (auto& c : cands) {
        es.eigenvalues.push_back(c.value);
        es.residuals.push_back(c.residual);
      }
      es.iterations = st.size;
      long passes = 0, seconds = 0;
      for (int j = 0; j < st.size; ++j) {
        const int hj = back_hist_[size_t(i) * m_ + j];
        passes += 1 + (hj ? 1 : 0);
        seconds += hj ? 1 : 0;
      }
      es.reorthogonalization_passes = passes;
      es.conditional_second_passes = seconds;
      es.converged = true;
      es.fp64_iterations = st.matvecs;
      if (want_vectors_) {
        const size_t n = seg_[i].n;
        for (auto& c : cands) {
          std::vector<cplx> v(n);
          QMBED_HIP(hipMemcpy(v.data(), Y + T_ * size_t(c.column) + size_t(seg_[i].tile0) * kTile, n * 16,
                              hipMemcpyDeviceToHost));
          es.eigenvectors.push_back(std::move(v));
        }
      }
    }
  }

  std::vector<Operator*> ops_;
  EigshOptions o_;
  EigshControls ctl_;
  int m_, ld_, k_;
  hipStream_t s_;
  int S_ = 0;
  uint32_t tiles_ = 0;
  size_t T_ = 0;
  std::vector<Seg> seg_;
  std::vector<uint32_t> tile_sector_;
  std::unique_ptr<SectorBatchAction> action_;
  DeviceArray<double> V_, w_, Y_, Z_, partial_, pa_, pn_, coef_, h_, w0n_, slot_, ritz_coef_;
  DeviceArray<int> hist_, sizes_;
  DeviceArray<uint32_t> list_, act_tiles_, act_secs_, act_hv_, r_tiles_, r_secs_, r_hv_, coef_off_, ts_d_;
  DeviceArray<uint8_t> seg_d_;
  uint32_t* stage_ = nullptr;
  size_t stage_words_ = 0;
  double* ritz_stage_ = nullptr;
  double* back_ = nullptr;
  int* back_hist_ = nullptr;
  double* slot_host_ = nullptr;
  hipEvent_t ev_ = nullptr;
  std::vector<SectorState> state_;
  std::vector<int> running_;
  int n_tiles_ = 0, n_secs_ = 0, n_hv_ = 0;
  long steps_ = 0;
  int ritz_rounds_ = 0;
  bool want_vectors_ = false;
};

}  // namespace

std::vector<Eigensystem> eigsh_sectors(const std::vector<Operator*>& ops, const EigshOptions& o,
                                       const EigshControls& ctl, bool want_vectors, hipStream_t s,
                                       SectorBatchStats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  if (o.target == SpectrumTarget::Shift) {
    // every sector runs its own shift-invert solve
    std::vector<Eigensystem> out;
    for (auto* op : ops) out.push_back(eigsh(*op, o, nullptr, ctl, want_vectors, s));
    if (stats) {
      stats->sectors = int(ops.size());
      stats->fallback = int(ops.size());
    }
    return out;
  }
  if (ctl.precision != MatvecPrecision::FP64) throw unsupported_backend("the sector batch runs FP64 matvecs");
  if (ops.size() > size_t(kMaxSectors)) throw unsupported_backend("sector batch: at most 4096 sectors");
  for (auto* op : ops) o.validate(op->dim());
  // Sectors for the lockstep batch: Hermitian Krylov-route sectors (the
  // dense route of small sectors and non-Hermitian sectors go one by one)
  std::vector<int> batched, single;
  long window = o.krylov_dimension ? std::min<long>(*o.krylov_dimension, o.max_iterations)
                                   : std::min<long>(std::max<long>(8L * o.eigenpairs + 64, 256), o.max_iterations);
  for (size_t i = 0; i < ops.size(); ++i) {
    const size_t n = ops[i]->dim();
    const bool dense = n <= 128 || (n <= 256 && !o.krylov_dimension);
    if (dense || !ops[i]->hermitian()) single.push_back(int(i));
    else {
      batched.push_back(int(i));
      window = std::min<long>(window, long(n));
    }
  }
  if (window > kMaxWindow) throw unsupported_backend("sector batch: the Krylov window exceeds 1024");
  if (!batched.empty() && window <= o.eigenpairs)
    throw invalid_options("the effective Krylov dimension must exceed eigenpairs");
  std::vector<Eigensystem> out(ops.size());
  SectorBatchStats st;
  st.sectors = int(ops.size());
  // Batches sized to the device: the Krylov basis of a batch (window + 2 + k
  // concatenated columns of 16 bytes per padded row, plus its partials) must
  // fit in 80% of the free memory; larger sets run as consecutive batches of
  // whole sectors, in input order.
  std::vector<std::vector<int>> chunks;
  if (!batched.empty()) {
    size_t free_b = 0, total_b = 0;
    QMBED_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per_row = (size_t(window) + 2 + size_t(o.eigenpairs)) * 16 + (size_t(window) + 1) * 16 / kTile;
    size_t budget = free_b / 10 * 8;
    if (ctl.batch_rows > 0) budget = std::min(budget, size_t(ctl.batch_rows) * per_row);
    std::vector<int> cur;
    size_t rows = 0;
    for (int i : batched) {
      const size_t padded = (ops[size_t(i)]->dim() + kTile - 1) / kTile * kTile;
      if (!cur.empty() && (rows + padded) * per_row > budget) {
        chunks.push_back(cur);
        cur.clear();
        rows = 0;
      }
      if (padded * per_row > budget)
        throw unsupported_backend("sector batch: the Krylov window of one sector does not fit in device memory");
      cur.push_back(i);
      rows += padded;
    }
    if (!cur.empty()) chunks.push_back(cur);
  }
  for (const auto& chunk : chunks) {
    std::vector<Operator*> bops;
    for (int i : chunk) bops.push_back(ops[size_t(i)]);
    Batch b(bops, o, ctl, int(window), s);
    b.run(want_vectors);
    st.steps += b.steps();
    st.ritz_rounds += b.ritz_rounds();
    st.window = int(window);
    ++st.batches;
    for (size_t q = 0; q < chunk.size(); ++q) {
      const auto& ss = b.state(int(q));
      if (ss.accepted) {
        out[size_t(chunk[q])] = ss.es;
        ++st.batched;
      } else {
        single.push_back(chunk[q]);
        ++st.fallback;
      }
    }
  }
  std::sort(single.begin(), single.end());
  for (int i : single) {
    DeviceArray<double> vecs;
    bool cv = false;
    Eigensystem es = eigsh_on_device(*ops[size_t(i)], o, nullptr, false, ctl, vecs, cv, s);
    if (want_vectors) {
      const size_t n = ops[size_t(i)]->dim(), e = cv ? 2 : 1;
      std::vector<double> h(e * n * es.eigenvalues.size());
      QMBED_HIP(hipMemcpy(h.data(), vecs.data(), h.size() * 8, hipMemcpyDeviceToHost));
      for (size_t q = 0; q < es.eigenvalues.size(); ++q) {
        std::vector<cplx> v(n);
        for (size_t r = 0; r < n; ++r)
          v[r] = cv ? cplx(h[e * (q * n + r)], h[e * (q * n + r) + 1]) : cplx(h[q * n + r], 0.0);
        es.eigenvectors.push_back(std::move(v));
      }
    }
    out[size_t(i)] = std::move(es);
  }
  QMBED_HIP(hipStreamSynchronize(s));
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  return out;
}

}  // namespace qmbed
