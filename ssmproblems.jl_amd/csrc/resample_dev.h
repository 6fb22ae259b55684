// resample_dev.h -- device code of the resamplers, shared by the single-filter kernels
// (resample.hip) and the batched serving kernels (serve.hip).
//
// Reference semantics (paths relative to the reference checkout):
//   Systematic (CPU)  GF/src/resamplers.jl:216-238  a_k = min{j : n*W_j > u0 + (k-1)}
//   Systematic (GPU)  GF/ext/CUDAExt.jl:49-58       O_j = min(n, floor(n*W_j/W_n + u0))
//   Stratified        GF/ext/CUDAExt.jl:68-78       O_j = min(n, floor(r_j + u[min(n, floor(r_j+1))]))
//   Multinomial       GF/ext/CUDAExt.jl:31-39       a_k = min{j : W_j >= u_k W_n}
#pragma once
#include <type_traits>

#include "device_common.h"
#include "internal.h"
#include "pf_common.h"

namespace ssm {

constexpr int kRsThreads = kTileThreads;

// The pointers a tile reads first lead the block, next to each other, so that the kernel's
// first argument load brings them together.
struct RsParams {
  int kind, n, mode;
  const void* w;
  const BfScalars* sc;
  const uint64_t* seed_dev = nullptr;  // when set, the Philox key is read from here
  double* prefix;
  const double* blk_m;
  const double* blk_c;
  const double* u;
  uint64_t seed;
  uint32_t step_host;
  double threshold_n;
  int32_t* anc;
  double* w_debug64;
  double* tile_sum;  // explicit modes: tile sums (the partials area is reused)
  double* Wbuf;   // multinomial: cumulative weights, tile-padded (nb * kTile)
  double* Wmid;   // multinomial: every 4th of them (nb * kTile / 4)
  double* Wroot;  // multinomial: every 64th of them (nb * kTile / 64)
  int32_t* counts_out;  // optional offspring count per particle (systematic, stratified)
};

__device__ __forceinline__ uint64_t rs_seed(const RsParams& p) {
  return p.seed_dev ? *p.seed_dev : p.seed;
}

__device__ __forceinline__ uint32_t rs_step(const RsParams& p) {
  return p.sc ? (uint32_t)p.sc->step : p.step_host;
}

// The scalars a resampling block decides with: the decision, the lazy-uniform flag (log-
// weight modes), the Philox step and key.  PLAIN (a particle filter's resample: the filter
// scalars and the device key are always there) reads them with no branch between the
// loads; its callers then wait for them together with the block's other inputs (an empty
// asm takes them all): a load behind a condition on a loaded value waits for the one
// before it, and the compiler otherwise moves loads behind the block's early return on the
// decision.
struct RsScalars {
  bool rs, typeless;
  uint32_t step;
  uint64_t seed;
  double ess;  // PLAIN: the loaded values, for the caller's wait
  int32_t typeless_in, step_in;
};
template <bool PLAIN>
__device__ __forceinline__ RsScalars rs_scalars(const RsParams& p, bool log_weights) {
  RsScalars r;
  if constexpr (PLAIN) {
    const double ess = p.sc->ess;
    const int32_t typeless = p.sc->typeless, step = p.sc->step;
    const uint64_t key = *p.seed_dev;
    r.ess = ess;
    r.typeless_in = typeless;
    r.step_in = step;
    r.rs = p.threshold_n >= ess;
    r.typeless = log_weights && typeless != 0;
    r.step = (uint32_t)step;
    r.seed = key;
  } else {
    r.rs = p.sc == nullptr || p.threshold_n >= p.sc->ess;
    r.typeless = log_weights && p.sc != nullptr && p.sc->typeless != 0;
    r.step = rs_step(p);
    r.seed = rs_seed(p);
  }
  return r;
}

// Weight of particle j in tile b as the resampler consumes it (FP64).  Particle-filter
// modes: softmax of the log-weights as w_j = exp(l_j - m_b) c_b (containers.jl:140-141),
// or 1/n for the lazy TypelessZero weights of initialise (containers.jl:62-65).
template <int MODE>
__device__ __forceinline__ double rs_weight(const RsParams& p, int j, int b) {
  if (MODE == kWeightsF64) return ((const double*)p.w)[j];
  if (MODE == kWeightsF32) return (double)((const float*)p.w)[j];
  if (p.sc->typeless) return 1.0 / (double)p.n;
  if (MODE == kWeightsLogF32) {
    const float e = exp_w(((const float*)p.w)[j] - (float)p.blk_m[b]);
    return (double)e * p.blk_c[b];
  } else {
    const double e = exp_w(((const double*)p.w)[j] - p.blk_m[b]);
    return e * p.blk_c[b];
  }
}

// The raw stored value behind a weight (loaded early), and the weight made from it with
// the tile's m_b and c_b: exactly what rs_weight returns.
template <int MODE>
struct RawW {
  using type = typename std::conditional<(MODE == kWeightsF64 || MODE == kWeightsLogF64), double,
                                         float>::type;
  type v;
};
template <int MODE>
__device__ __forceinline__ RawW<MODE> rs_raw(const RsParams& p, int j) {
  using V = typename RawW<MODE>::type;
  return RawW<MODE>{((const V*)p.w)[j]};
}
template <int MODE>
__device__ __forceinline__ double rs_weight_of(const RsParams& p, RawW<MODE> r, double m_b,
                                               double c_b, bool typeless) {
  if (MODE == kWeightsF64 || MODE == kWeightsF32) return (double)r.v;
  if (typeless) return 1.0 / (double)p.n;
  if (MODE == kWeightsLogF32) {
    const float e = exp_w((float)r.v - (float)m_b);
    return (double)e * c_b;
  } else {
    const double e = exp_w((double)r.v - m_b);
    return e * c_b;
  }
}

// Inclusive FP64 scan across the block (thread order), fixed order; *tot = block total.
__device__ __forceinline__ double block_incl_scan_d(double v, double* lds, double* tot) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const double incl = wave_scan_dpp(v);
  __syncthreads();
  if (lane == kWave - 1) lds[w] = incl;
  __syncthreads();
  double wp = 0.0;
  for (int k = 0; k < w; ++k) wp += lds[k];
  *tot = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  return wp + incl;
}

// Inclusive max-scan across the block (thread order); exact.  *tot = block max.
__device__ __forceinline__ double block_incl_max_d(double v, double* lds, double* tot) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  v = wave_scan_max_dpp(v);
  __syncthreads();
  if (lane == kWave - 1) lds[w] = v;
  __syncthreads();
  for (int k = 0; k < w; ++k) v = fmax(v, lds[k]);
  *tot = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
  return v;
}

// Cumulative offspring count for an inclusive prefix W.
// SYSTEMATIC: O = #{s in [0,n) : fl(u0 + s) < fl(n W)} (strict, resamplers.jl:231).
// The estimate ceil(fl(n W - u0)) is off by at most one: the rounding of n W - u0 and of
// u0 + s both stay below 2^-28 for n < 2^31 (|u0| < 1), so the two can only disagree when
// n W - u0 lies that close to an integer K, and then the estimate and the count are both K
// or K + 1.  One exact test in each direction therefore settles it, branch-free.
__device__ __forceinline__ int offspring_systematic(double W, double u0, int n) {
  const double vs = W * (double)n;
  double est = ceil(vs - u0);
  est = (est >= 0.0) ? est : 0.0;  // also catches NaN
  est = fmin(est, (double)n);
  const int m = (int)est;
  const bool down = (m > 0) && !(u0 + (double)(m - 1) < vs);
  const bool up = (m < n) && (u0 + (double)m < vs);
  return m + (int)up - (int)down;
}

__device__ __forceinline__ int clamp_floor(double v, int n) {
  double f = floor(v);
  if (!(f >= 0.0)) f = 0.0;
  if (f > (double)n) f = (double)n;
  return (int)f;
}

// KIND >= 0 fixes the resampler at compile time (-1: p.kind); PLAIN: the uniforms come from
// Philox (no explicit uniform array), so the specialised kernels have no loads under
// conditions.
template <int KIND = -1, bool PLAIN = false>
__device__ __forceinline__ int offspring_of(const RsParams& p, double W, double Wn, double u0,
                                            uint32_t step, uint64_t seed) {
  const int n = p.n;
  const int kind = KIND >= 0 ? KIND : p.kind;
  if (kind == SSM_SYSTEMATIC) return offspring_systematic(W, u0, n);
  const double r = ((double)n * W) / Wn;
  if (kind == SSM_SYSTEMATIC_MURRAY) return clamp_floor(r + u0, n);
  // STRATIFIED: k = min(n, floor(r + 1)) is 1-based into u.
  int k = clamp_floor(r + 1.0, n);
  if (k < 1) k = 1;
  const double* u = PLAIN ? nullptr : p.u;
  const double uk = u ? u[k - 1] : resample_uniform(seed, step, (uint32_t)(k - 1));
  return clamp_floor(r + uk, n);
}

// Output slots one wave fills per round of its ancestor fill (4 per lane).
constexpr int kFillWin = kWave * kTileItems;
// Multinomial search index: per tile, the cumulative weight closing every group of 64
// positions (16 per tile, one 128-byte node) and every group of 4 (16 per 64-group).
constexpr int kRootGroup = 64, kRootPerTile = kTile / kRootGroup;
constexpr int kMidGroup = 4, kMidPerTile = kTile / kMidGroup;
// Entries of the LDS-resident top level of the multinomial search.
constexpr int kCoarse = 1024;

// What a tile's resample needs besides its weights: the decision, the lazy-uniform flag,
// the tile's prefix P, the next tile's prefix, the total, the tile's weight scale (m_b, c_b)
// and the Philox step of the uniforms.
struct TileIn {
  bool rs, typeless;
  double P, Pnext, Wn, m_b, c_b;
  uint32_t step;
  uint64_t seed;
};

// Loads the raw weights of tile b for thread t's positions 4t..4t+3 (clamped in-bounds
// indices; a full tile is read with one vector load per thread).  WHOLE: the weights hold
// whole tiles (a particle filter's Npad entries) in a 16-byte aligned buffer, so every
// tile is read with the vector loads and no branch (positions past n are masked by the
// caller).
template <int MODE, bool WHOLE = false>
__device__ __forceinline__ void load_tile_raw(const RsParams& p, int b,
                                              RawW<MODE> raw[kTileItems]) {
  using V = typename RawW<MODE>::type;
  const int n = p.n, tile0 = b * kTile;
  const int j0 = tile0 + kTileItems * threadIdx.x;
  const V* w = (const V*)p.w;
  if (WHOLE || (tile0 + kTile <= n && ((uintptr_t)w & 15) == 0)) {
    if constexpr (sizeof(V) == 4) {
      const float4 v = *reinterpret_cast<const float4*>(w + j0);
      raw[0].v = v.x;
      raw[1].v = v.y;
      raw[2].v = v.z;
      raw[3].v = v.w;
    } else {
      const double2 a = *reinterpret_cast<const double2*>(w + j0);
      const double2 c = *reinterpret_cast<const double2*>(w + j0 + 2);
      raw[0].v = a.x;
      raw[1].v = a.y;
      raw[2].v = c.x;
      raw[3].v = c.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) raw[k].v = w[min(j0 + k, n - 1)];
  }
}

// The cumulative weights of tile b's particles, W[k] for position 4t + k: the tile's prefix
// P plus the in-tile sums, nondecreasing in position by construction.  Thread t sums its 4
// consecutive weights serially (acc_k); a wave scan of the thread sums gives each thread's
// in-wave start x_t, and y_k = x_t + acc_k is the in-wave cumulative weight, made
// nondecreasing across lanes by an exact running max.  Lane 63 of each wave publishes the
// wave total and the wave's largest y; after the one barrier every thread forms the wave
// offsets off_{v+1} = max(off_v + tot_v, off_v + ymax_v), so no value of wave v exceeds the
// first value of wave v + 1, and W = P + (off_w + y).  Positions past n add zero.  Also
// returns the tile's prefix, the next tile's prefix, the total (for the lazy TypelessZero
// weights: tile0/n, the next tile's start over n or 1, and 1) and *Wedge, the W of the last
// position of the previous wave (of the tile's first wave: P).  Optionally records each
// consumed FP64 weight in w_debug64.
template <int MODE>
__device__ __forceinline__ void tile_cum_weights(const RsParams& p, int b,
                                                 const RawW<MODE> raw[kTileItems],
                                                 const TileIn& in, double* w_debug64,
                                                 double W[kTileItems], double* P_out,
                                                 double* Pnext_out, double* Wn_out,
                                                 double* Wedge_out) {
  constexpr int NW = kRsThreads / kWave;
  __shared__ double lds_tot[NW], lds_ymax[NW];
  const int n = p.n;
  const int nb = (n + kTile - 1) / kTile;
  const int tile0 = b * kTile;
  const int t = threadIdx.x, w = t / kWave, lane = t & (kWave - 1);
  double P, Pnext, Wn;
  if (in.typeless) {
    P = (double)tile0 / (double)n;
    Pnext = (b == nb - 1) ? 1.0 : (double)(tile0 + kTile) / (double)n;
    Wn = 1.0;
  } else {
    P = in.P;
    Pnext = in.Pnext;
    Wn = in.Wn;
  }
  double acc[kTileItems];
  double run = 0.0;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int j = tile0 + kTileItems * t + k;
    // computed for every position (raw holds clamped in-bounds values), masked past n
    double wj = rs_weight_of<MODE>(p, raw[k], in.m_b, in.c_b, in.typeless);
    wj = (j < n) ? wj : 0.0;
    if (w_debug64 && j < n) w_debug64[j] = wj;
    run += wj;
    acc[k] = run;
  }
  const double incl = wave_scan_dpp(run);
  const double x = dpp_d<kDppWaveShr1>(0.0, incl);  // the previous lane's; lane 0: 0
  double y[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) y[k] = x + acc[k];  // nondecreasing in k
  const double ym_incl = wave_scan_max_dpp(y[kTileItems - 1]);
  const double ym_prev = dpp_d<kDppWaveShr1>(0.0, ym_incl);
  if (lane == kWave - 1) {
    lds_tot[w] = incl;
    lds_ymax[w] = ym_incl;
  }
  __syncthreads();
  double off = 0.0, edge = P;
  for (int v = 0; v < w; ++v) {
    edge = P + (off + lds_ymax[v]);
    off = fmax(off + lds_tot[v], off + lds_ymax[v]);
  }
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) W[k] = P + (off + fmax(y[k], ym_prev));
  *P_out = P;
  *Pnext_out = Pnext;
  *Wn_out = Wn;
  *Wedge_out = edge;
}

// One wave's ancestor fill.  The wave's particles own the contiguous output slots
// [wstart, wend); particle k of this lane owns [start[k], O[k]).  Slots are written in
// windows of kFillWin: every particle with offspring marks the slot where its range starts
// with its position in the wave; lane l then takes the window's slots 4l..4l+3, and a
// running max over them (serial in the lane, then an exclusive max scan across the lanes,
// then the carry of the previous window) assigns every slot the last particle that started
// at or before it.  The results go back through LDS so each store instruction writes 64
// consecutive slots.  A particle whose range covers a whole window is written directly
// (ow holds every position's range end), which keeps degenerate weights balanced.  anc0 is
// the particle index of the wave's position 0.
template <bool WT>
__device__ __forceinline__ void wave_fill(int* marks, const int* ow,
                                          const int start[kTileItems],
                                          const int O[kTileItems], int wstart, int wend,
                                          int anc0, int32_t* __restrict__ anc) {
  const int lane = threadIdx.x & (kWave - 1);
  int4* const mine = reinterpret_cast<int4*>(marks) + lane;
  int base = wstart, carry = -1;
  while (base < wend) {
    if (carry >= 0) {
      const int e = __builtin_amdgcn_readfirstlane(min(ow[carry], wend));
      if (e - base >= kFillWin) {
        const int v = anc0 + carry;
        for (int s = base + lane; s < e; s += kWave) store_out<WT>(&anc[s], v);
        base = e;
        continue;
      }
    }
    *mine = make_int4(-1, -1, -1, -1);
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      const int s = start[k] - base;
      if (O[k] > start[k] && s >= 0 && s < kFillWin) marks[s] = kTileItems * lane + k;
    }
    wave_lds_sync();
    int4 m = *mine;
    m.y = max(m.y, m.x);
    m.z = max(m.z, m.y);
    m.w = max(m.w, m.z);
    // the previous lane's running max (lane 0: none), then the earlier windows' carry
    const int before = max(dpp_i<kDppWaveShr1>(-1, wave_scan_max_i32(m.w)), carry);
    m.x = max(m.x, before);
    m.y = max(m.y, before);
    m.z = max(m.z, before);
    m.w = max(m.w, before);
    carry = __builtin_amdgcn_readlane(m.w, kWave - 1);
    wave_lds_sync();
    *mine = m;
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < kTileItems; ++i) {
      const int s = base + i * kWave + lane;
      const int a = marks[i * kWave + lane];
      if (s < wend) store_out<WT>(&anc[s], anc0 + a);
    }
    wave_lds_sync();
    base += kFillWin;
  }
}

// Tile rescan -> cumulative offspring counts -> ancestor fill, for one tile.  KIND/PLAIN as
// offspring_of; PLAIN also means no debug weights and no offspring counts.  Thread t owns
// the 4 consecutive particles 4t..4t+3 of the tile.  Each particle's cumulative offspring
// count O comes from its cumulative weight; the wave's output range starts at the count of
// the previous wave's last position (recomputed from its published cumulative weight, so
// no second barrier), and each wave writes its own range (wave_fill).  For MULTINOMIAL it
// writes the nondecreasing cumulative weights of every position of the tile to Wbuf
// (positions past n hold the tile's closing prefix), every 4th of them to Wmid and every
// 64th to Wroot.
// All threads of the block call it; it returns at once when !in.rs.  WT: the ancestors
// are stored at device scope (a single filter's launch); the server's batched resampler
// stores them plainly.
template <int MODE, int KIND = -1, bool PLAIN = false, bool WT = true>
__device__ __forceinline__ void scan_offspring_tile_in(const RsParams& p, int b,
                                                       const RawW<MODE> raw[kTileItems],
                                                       const TileIn& in) {
  if (!in.rs) return;
  constexpr int NW = kRsThreads / kWave;
  const int kind = KIND >= 0 ? KIND : p.kind;
  double* const w_debug64 = PLAIN ? nullptr : p.w_debug64;
  int32_t* const counts_out = PLAIN ? nullptr : p.counts_out;
  const double* const uvec = PLAIN ? nullptr : p.u;
  __shared__ __align__(16) int o_lds[kTile];
  __shared__ __align__(16) int marks[NW][kFillWin];
  const int n = p.n;
  const int nb = (n + kTile - 1) / kTile;
  const int tile0 = b * kTile;
  const int t = threadIdx.x, w = t / kWave, lane = t & (kWave - 1);
  const int cnt = min(kTile, n - tile0);
  double W[kTileItems], P, Pnext, Wn, Wedge;
  tile_cum_weights<MODE>(p, b, raw, in, w_debug64, W, &P, &Pnext, &Wn, &Wedge);
  const uint32_t step = in.step;
  if (kind == SSM_MULTINOMIAL) {
    double v[kTileItems];
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      const int pos = kTileItems * t + k;
      v[k] = (pos >= cnt - 1) ? Pnext : fmin(W[k], Pnext);
    }
    double2* dst = reinterpret_cast<double2*>(p.Wbuf + (size_t)tile0 + kTileItems * t);
    dst[0] = make_double2(v[0], v[1]);
    dst[1] = make_double2(v[2], v[3]);
    p.Wmid[(size_t)b * kMidPerTile + t] = v[kTileItems - 1];
    if ((t & 15) == 15) p.Wroot[(size_t)b * kRootPerTile + (t >> 4)] = v[kTileItems - 1];
    return;
  }
  const double u0 = (kind == SSM_STRATIFIED)
                        ? 0.0
                        : (uvec ? uvec[0] : resample_uniform(in.seed, step, 0u));
  // The wave-uniform counts (the tile's start and end, the wave's first slot) are one
  // lane's work each and then read from it: four waves share a SIMD's FP64 issue.
  const double Wq = (lane == 0) ? P : (lane == 1) ? Pnext : Wedge;
  const int oq = offspring_of<KIND, PLAIN>(p, Wq, Wn, u0, step, in.seed);
  const int o_prev = (b == 0) ? 0 : __builtin_amdgcn_readlane(oq, 0);
  const int o_end = (b == nb - 1) ? n : __builtin_amdgcn_readlane(oq, 1);
  const int o_edge = __builtin_amdgcn_readlane(oq, 2);
  // the tile's last particle closes its range; positions past n own none
  auto count_at = [&](int pos, double Wv) {
    return (pos >= cnt - 1) ? o_end
                            : min(max(offspring_of<KIND, PLAIN>(p, Wv, Wn, u0, step, in.seed),
                                      o_prev),
                                  o_end);
  };
  int O[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) O[k] = count_at(kTileItems * t + k, W[k]);
  const int wstart = (w == 0)                       ? o_prev
                     : (kFillWin * w - 1 >= cnt - 1) ? o_end
                                                     : min(max(o_edge, o_prev), o_end);
  const int s0 = dpp_i<kDppWaveShr1>(wstart, O[kTileItems - 1]);  // lane 0: wstart
  const int start[kTileItems] = {s0, O[0], O[1], O[2]};
  if (counts_out) {  // particle k owns the output slots [start[k], O[k])
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      const int pos = kTileItems * t + k;
      if (pos < cnt) counts_out[tile0 + pos] = O[k] - start[k];
    }
  }
  int* ow = o_lds + w * kFillWin;
  *reinterpret_cast<int4*>(ow + kTileItems * lane) = make_int4(O[0], O[1], O[2], O[3]);
  wave_lds_sync();
  const int wend = __builtin_amdgcn_readlane(O[kTileItems - 1], kWave - 1);
  wave_fill<WT>(marks[w], ow, start, O, wstart, wend, tile0 + w * kFillWin, p.anc);
}

// The resample of tile b with its inputs from memory: the decision and flags from the
// filter's scalars (or always, for explicit weights), the prefixes and scales from the
// finish kernel.  Every global input is requested before the decision is known (all the
// addresses are in bounds), so the kernel pays one round trip before its arithmetic.
template <int MODE, int KIND = -1, bool PLAIN = false, bool WT = true>
__device__ __forceinline__ void scan_offspring_tile(const RsParams& p, int b) {
  const int nb = (p.n + kTile - 1) / kTile;
  RawW<MODE> raw[kTileItems];
  load_tile_raw<MODE, PLAIN>(p, b, raw);
  constexpr bool kLog = (MODE == kWeightsLogF32 || MODE == kWeightsLogF64);
  TileIn in;
  in.P = p.prefix[b];
  in.Pnext = p.prefix[b + 1];
  in.Wn = p.prefix[nb];
  in.m_b = kLog ? p.blk_m[b] : 0.0;
  in.c_b = kLog ? p.blk_c[b] : 0.0;
  const RsScalars rsc = rs_scalars<PLAIN>(p, kLog);
  if constexpr (PLAIN) {
    // one wait for the scalars, the tile's prefixes and scales and its weights
    asm volatile("" ::"s"(rsc.ess), "s"(rsc.typeless_in), "s"(rsc.step_in), "s"(rsc.seed),
                 "s"(in.P), "s"(in.Pnext), "s"(in.Wn), "s"(in.m_b), "s"(in.c_b));
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) asm volatile("" : "+v"(raw[k].v));
  }
  in.rs = rsc.rs;
  in.typeless = rsc.typeless;
  in.step = rsc.step;
  in.seed = rsc.seed;
  scan_offspring_tile_in<MODE, KIND, PLAIN, WT>(p, b, raw, in);
}

// The cumulative weight that closes tile b: the last entry of its Wbuf positions.
__device__ __forceinline__ double tile_close(const RsParams& p, int b, int nb, bool typeless,
                                             double Wn) {
  if (b >= nb - 1) return Wn;
  return typeless ? (double)((b + 1) * kTile) / (double)p.n : p.prefix[b + 1];
}

// The number of the first L - 1 of L nondecreasing values below q (the lower bound of q when
// the last value is at least q): one vector load of the node, no dependent steps.
template <int L>
__device__ __forceinline__ int node_rank(const double* __restrict__ node, double q) {
  const double2* v2 = reinterpret_cast<const double2*>(node);
  double2 v[L / 2];
#pragma unroll
  for (int i = 0; i < L / 2; ++i) v[i] = v2[i];
  int r = 0;
#pragma unroll
  for (int i = 0; i < L / 2; ++i) {
    r += (v[i].x < q) ? 1 : 0;
    if (2 * i + 1 < L - 1) r += (v[i].y < q) ? 1 : 0;
  }
  return r;
}

// Sum of v over each aligned team of 8 lanes, in every lane of the team: quad exchanges
// (xor 1, xor 2), then the other quad of the half-row (row_half_mirror).
__device__ __forceinline__ int team8_sum(int v) {
  v += dpp_i<0xB1>(0, v);   // quad_perm [1, 0, 3, 2]
  v += dpp_i<0x4E>(0, v);   // quad_perm [2, 3, 0, 1]
  v += dpp_i<0x141>(0, v);  // row_half_mirror
  return v;
}

// node_rank for a team of 8 lanes: lane tl holds entries 2 tl and 2 tl + 1 of the node (one
// 16-byte load; the team's loads together are the node's line), counts those of the first
// L - 1 below q, and the team sums the counts.
template <int L>
__device__ __forceinline__ int team_node_rank(const double* __restrict__ node, double q,
                                              int tl) {
  static_assert(L == 16 || L == 4, "16- and 4-entry nodes");
  int c = 0;
  if (2 * tl < L) {
    const double2 v = reinterpret_cast<const double2*>(node)[tl];
    c = ((v.x < q) ? 1 : 0) + ((2 * tl + 1 < L - 1 && v.y < q) ? 1 : 0);
  }
  return team8_sum(c);
}

// Multinomial: a_k = min{j : W_j >= u_k * W_n} for the queries k of block `blk`
// (blk * kTile + j, j < 1024), with W the nondecreasing cumulative weights the scan wrote
// to Wbuf.  Each lane first finds the tiles of 4 queries by a branch-free lower-bound
// search over the tiles' closing prefixes (its top level -- every 2^sh-th close, at most
// kCoarse of them -- sits in LDS) and leaves (q, tile) in LDS for its wave.  Within the
// tile a three-level tree of 16-way nodes (Wroot: the closes of the 64-position groups,
// Wmid: of the 4-position groups, then Wbuf) needs three dependent node loads; each node's
// last entry is at least the query, so the result equals a search of the whole Wbuf.  The
// queries are random, so in a one-lane-per-query search every lane of a load touches its
// own cache line and the texture units' per-line lookups bound the kernel; here a team of
// 8 lanes walks the tree for one query, each lane loading 16 bytes of the node, so a
// wave's load instruction touches 8 lines instead of 64.  A wave's 256 queries go in
// rounds of kTeamBatch per team, their loads in flight together.  All threads of the
// block call it.
constexpr int kTeamBatch = 8;
template <bool PLAIN>
__device__ __forceinline__ void multinomial_search_block(const RsParams& p, int blk) {
  __shared__ double coarse[kCoarse];
  __shared__ double qs[kTile];
  __shared__ int tiles[kTile];
  const int n = p.n, nb = (n + kTile - 1) / kTile, t = threadIdx.x;
  const RsScalars rsc = rs_scalars<PLAIN>(p, p.mode == kWeightsLogF32 || p.mode == kWeightsLogF64);
  const double Wtot = p.prefix[nb];  // read before the decision (prefix holds nb + 1 entries)
  if constexpr (PLAIN)
    asm volatile("" ::"s"(rsc.ess), "s"(rsc.typeless_in), "s"(rsc.step_in), "s"(rsc.seed),
                 "s"(Wtot));
  if (!rsc.rs) return;
  const uint32_t step = rsc.step;
  const bool typeless = rsc.typeless;
  const double Wn = typeless ? 1.0 : Wtot;
  int sh = 0;
  while (((nb + (1 << sh) - 1) >> sh) > kCoarse) ++sh;
  const int ns = (nb + (1 << sh) - 1) >> sh;
  for (int i = t; i < ns; i += kRsThreads)
    coarse[i] = tile_close(p, min((i + 1) << sh, nb) - 1, nb, typeless, Wn);
  __syncthreads();
  const int w = t / kWave, lane = t & (kWave - 1), tm = lane >> 3, tl = lane & 7;
  constexpr int kPerWave = kTile / (kRsThreads / kWave), kTeams = kWave / 8;
  // the tiles of this wave's queries j = 256 w + 64 i + lane, two in flight at a time
#pragma unroll
  for (int i0 = 0; i0 < kTileItems; i0 += 2) {
#pragma unroll
    for (int i = i0; i < i0 + 2; ++i) {
      const int j = w * kPerWave + i * kWave + lane;
      const int kc = min(blk * kTile + j, n - 1);
      const double uk = p.u ? p.u[kc] : resample_uniform(rsc.seed, step, (uint32_t)kc);
      const double q = uk * Wn;
      int c = 0;
#pragma unroll
      for (int h = kCoarse / 2; h >= 1; h >>= 1)
        if (c + h <= ns - 1 && coarse[c + h - 1] < q) c += h;
      const int g0 = c << sh, gl = min(g0 + (1 << sh), nb) - g0;
      int bt = 0;
      for (int h = (1 << sh) >> 1; h >= 1; h >>= 1)
        if (bt + h <= gl - 1 && tile_close(p, g0 + bt + h - 1, nb, typeless, Wn) < q) bt += h;
      qs[j] = q;
      tiles[j] = g0 + bt;
    }
  }
  wave_lds_sync();  // each wave reads back only its own queries
  for (int r0 = 0; r0 < kPerWave / kTeams; r0 += kTeamBatch) {
    int jj[kTeamBatch], bs[kTeamBatch], g1[kTeamBatch], g2[kTeamBatch];
    double q[kTeamBatch];
#pragma unroll
    for (int b = 0; b < kTeamBatch; ++b) {
      jj[b] = w * kPerWave + (r0 + b) * kTeams + tm;
      q[b] = qs[jj[b]];
      bs[b] = tiles[jj[b]];
    }
#pragma unroll
    for (int b = 0; b < kTeamBatch; ++b)
      g1[b] = team_node_rank<kRootPerTile>(p.Wroot + (size_t)bs[b] * kRootPerTile, q[b], tl);
#pragma unroll
    for (int b = 0; b < kTeamBatch; ++b)
      g2[b] = team_node_rank<kRootGroup / kMidGroup>(
          p.Wmid + (size_t)bs[b] * kMidPerTile + g1[b] * (kRootGroup / kMidGroup), q[b], tl);
#pragma unroll
    for (int b = 0; b < kTeamBatch; ++b) {
      const int e = team_node_rank<kMidGroup>(
          p.Wbuf + (size_t)bs[b] * kTile + g1[b] * kRootGroup + g2[b] * kMidGroup, q[b], tl);
      const int k = blk * kTile + jj[b];
      if (tl == 0 && k < n)
        p.anc[k] = min(bs[b] * kTile + g1[b] * kRootGroup + g2[b] * kMidGroup + e, n - 1);
    }
  }
}

}  // namespace ssm

namespace ssm {

// One categorical draw from the weights a finish kernel has folded (StatsBase.sample with
// Weights, as the multinomial rule of CUDAExt.jl:31-39 for a single uniform): the first
// particle j whose cumulative weight W_j reaches u W_n, with the same cumulative weights the
// multinomial resampler forms (per-tile prefixes from the finish, the in-tile scan of
// tile_cum_weights, the tile's last particle closing at the next prefix).  One block of
// kRsThreads threads; *out gets the index.  The uniform is u, or (u < 0) Philox draw k = 0
// of `stream` at the filter's current step.
template <int MODE>
__device__ __forceinline__ void categorical_one_block(const RsParams& p, double u,
                                                      uint32_t stream, int32_t* out) {
  __shared__ int cnt_lds[kRsThreads / kWave];
  __shared__ int tile_sel;
  const int n = p.n, nb = (n + kTile - 1) / kTile, t = threadIdx.x;
  constexpr bool kLog = (MODE == kWeightsLogF32 || MODE == kWeightsLogF64);
  const bool typeless = kLog && p.sc->typeless;
  const uint32_t step = rs_step(p);
  const double uu = (u >= 0.0) ? u : stream_uniform(rs_seed(p), stream, step, 0u);
  const double Wn = typeless ? 1.0 : p.prefix[nb];
  const double q = uu * Wn;
  // the tile: the number of tiles whose closing prefix is below q (tile b closes at
  // prefix[b + 1]; the last tile at Wn), counted over all tiles in parallel
  int below = 0;
  for (int b = t; b < nb; b += kRsThreads) {
    const double close = (b == nb - 1)
                             ? Wn
                             : (typeless ? (double)((b + 1) * kTile) / (double)n : p.prefix[b + 1]);
    below += (close < q) ? 1 : 0;
  }
  below = (int)wave_sum_dpp((float)below);  // at most nb < 2^24: exact
  const int w = t / kWave, lane = t & (kWave - 1);
  if (lane == 0) cnt_lds[w] = below;
  __syncthreads();
  if (t == 0) {
    int s = 0;
    for (int k = 0; k < kRsThreads / kWave; ++k) s += cnt_lds[k];
    tile_sel = min(s, nb - 1);
  }
  __syncthreads();
  const int b = tile_sel;
  RawW<MODE> raw[kTileItems];
  load_tile_raw<MODE>(p, b, raw);
  TileIn in;
  in.rs = true;
  in.typeless = typeless;
  in.P = p.prefix[b];
  in.Pnext = p.prefix[b + 1];
  in.Wn = Wn;
  in.m_b = kLog ? p.blk_m[b] : 0.0;
  in.c_b = kLog ? p.blk_c[b] : 0.0;
  in.step = step;
  in.seed = 0;  // the tile's weights only
  double W[kTileItems], P, Pnext, Wt, Wedge;
  tile_cum_weights<MODE>(p, b, raw, in, nullptr, W, &P, &Pnext, &Wt, &Wedge);
  const int tile0 = b * kTile, cnt = min(kTile, n - tile0);
  // the first position of the tile whose cumulative weight reaches q (the tile's last
  // particle closes at Pnext, as in the multinomial resampler's W array)
  int first = kTile;
#pragma unroll
  for (int k = kTileItems - 1; k >= 0; --k) {
    const int pos = kTileItems * t + k;
    const double Wk = (pos == cnt - 1) ? Pnext : fmin(W[k], Pnext);
    if (pos < cnt && Wk >= q) first = pos;
  }
  int m = first;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) m = min(m, __shfl_xor(m, off, kWave));
  if (lane == 0) cnt_lds[w] = m;
  __syncthreads();
  if (t == 0) {
    int s = kTile;
    for (int k = 0; k < kRsThreads / kWave; ++k) s = min(s, cnt_lds[k]);
    *out = tile0 + min(s, cnt - 1);
  }
}

}  // namespace ssm
