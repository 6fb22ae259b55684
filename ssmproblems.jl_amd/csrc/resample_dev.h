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

struct RsParams {
  int kind, n, mode;
  const void* w;
  const double* u;
  uint64_t seed;
  const uint64_t* seed_dev = nullptr;  // when set, the Philox key is read from here
  uint32_t step_host;
  const BfScalars* sc;
  double threshold_n;
  int32_t* anc;
  double* w_debug64;
  const double* blk_m;
  const double* blk_c;
  double* tile_sum;  // explicit modes: tile sums (the partials area is reused)
  double* prefix;
  double* Wbuf;
  int32_t* counts_out;  // optional offspring count per particle (systematic, stratified)
};

__device__ __forceinline__ bool rs_active(const RsParams& p) {
  return p.sc == nullptr || p.threshold_n >= p.sc->ess;
}

__device__ __forceinline__ uint64_t rs_seed(const RsParams& p) {
  return p.seed_dev ? *p.seed_dev : p.seed;
}

__device__ __forceinline__ uint32_t rs_step(const RsParams& p) {
  return p.sc ? (uint32_t)p.sc->step : p.step_host;
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
__device__ __forceinline__ int offspring_systematic(double W, double u0, int n) {
  const double vs = W * (double)n;
  double est = ceil(vs - u0);
  if (!(est >= 0.0)) est = 0.0;  // also catches NaN
  if (est > (double)n) est = (double)n;
  int m = (int)est;
  while (m > 0 && !(u0 + (double)(m - 1) < vs)) --m;
  while (m < n && (u0 + (double)m < vs)) ++m;
  return m;
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
                                            uint32_t step) {
  const int n = p.n;
  const int kind = KIND >= 0 ? KIND : p.kind;
  if (kind == SSM_SYSTEMATIC) return offspring_systematic(W, u0, n);
  const double r = ((double)n * W) / Wn;
  if (kind == SSM_SYSTEMATIC_MURRAY) return clamp_floor(r + u0, n);
  // STRATIFIED: k = min(n, floor(r + 1)) is 1-based into u.
  int k = clamp_floor(r + 1.0, n);
  if (k < 1) k = 1;
  const double* u = PLAIN ? nullptr : p.u;
  const double uk = u ? u[k - 1] : resample_uniform(rs_seed(p), step, (uint32_t)(k - 1));
  return clamp_floor(r + uk, n);
}

// Exclusive FP64 scan of one value per thread across the block (thread order, fixed
// order); *tot = block total.
__device__ __forceinline__ double block_excl_scan_d(double v, double* lds, double* tot) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const double incl = wave_scan_dpp(v);
  double ex = wave_shfl_up(incl, 1);
  if (lane == 0) ex = 0.0;
  __syncthreads();
  if (lane == kWave - 1) lds[w] = incl;
  __syncthreads();
  for (int k = 0; k < w; ++k) ex += lds[k];
  *tot = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  return ex;
}

// Exclusive max-scan of one value per thread across the block (exact).
__device__ __forceinline__ double block_excl_max_d(double v, double* lds) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  v = wave_scan_max_dpp(v);
  double ex = wave_shfl_up(v, 1);
  if (lane == 0) ex = -INFINITY;
  __syncthreads();
  if (lane == kWave - 1) lds[w] = v;
  __syncthreads();
  for (int k = 0; k < w; ++k) ex = fmax(ex, lds[k]);
  return ex;
}

// Slots a thread writes itself; threads with more hand their range to the whole block.
constexpr int kDirectSlots = 32;

// Main kernel: tile rescan -> cumulative offspring counts -> ancestor fill.  Thread t owns
// the 4 consecutive particles 4t..4t+3 of the tile, so the tile scan is a thread-serial
// scan plus one block scan.  Every particle's output slots [O_{j-1}, O_j) are written by
// its own thread (neighbouring threads write neighbouring ranges); a thread whose range
// exceeds kDirectSlots hands it to the whole block, which keeps degenerate weights (one
// particle taking every slot) balanced.  For MULTINOMIAL the kernel writes the
// nondecreasing prefix sums W to Wbuf instead.
// What a tile's resample needs besides its weights: the decision, the lazy-uniform flag,
// the tile's prefix P, the next tile's prefix, the total, the tile's weight scale (m_b, c_b)
// and the Philox step of the uniforms.
struct TileIn {
  bool rs, typeless;
  double P, Pnext, Wn, m_b, c_b;
  uint32_t step;
};

// Loads the raw weights of tile b (clamped in-bounds indices).
template <int MODE>
__device__ __forceinline__ void load_tile_raw(const RsParams& p, int b,
                                              RawW<MODE> raw[kTileItems]) {
  const int n = p.n, tile0 = b * kTile;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int j = tile0 + kTileItems * threadIdx.x + k;
    raw[k] = rs_raw<MODE>(p, j < n ? j : n - 1);
  }
}

// The cumulative weights of tile b's particles, W[k] for position kTileItems t + k: the
// tile's prefix P plus the in-tile sums (a thread-serial scan of kTileItems consecutive
// weights, then one block scan of the thread sums), made nondecreasing by an exact running
// max.  Positions past n add zero.  Also returns the tile's prefix, the next tile's prefix
// and the total (for the lazy TypelessZero weights: tile0/n, the next tile's start over n
// or 1, and 1).  Optionally records each consumed FP64 weight in w_debug64.
template <int MODE>
__device__ __forceinline__ void tile_cum_weights(const RsParams& p, int b,
                                                 const RawW<MODE> raw[kTileItems],
                                                 const TileIn& in, double* w_debug64,
                                                 double W[kTileItems], double* P_out,
                                                 double* Pnext_out, double* Wn_out) {
  __shared__ double lds[kRsThreads / kWave];
  const int n = p.n;
  const int nb = (n + kTile - 1) / kTile;
  const int tile0 = b * kTile;
  const int t = threadIdx.x;
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
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int j = tile0 + kTileItems * t + k;
    // computed for every position (raw holds clamped in-bounds values), masked past n
    double wj = rs_weight_of<MODE>(p, raw[k], in.m_b, in.c_b, in.typeless);
    wj = (j < n) ? wj : 0.0;
    if (w_debug64 && j < n) w_debug64[j] = wj;
    acc += wj;
    W[k] = acc;
  }
  double tot;
  const double ex = block_excl_scan_d(acc, lds, &tot);
  (void)tot;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) W[k] = P + (ex + W[k]);
  // nondecreasing in particle order (exact running max)
#pragma unroll
  for (int k = 1; k < kTileItems; ++k) W[k] = fmax(W[k], W[k - 1]);
  const double carry = block_excl_max_d(W[kTileItems - 1], lds);
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) W[k] = fmax(W[k], carry);
  *P_out = P;
  *Pnext_out = Pnext;
  *Wn_out = Wn;
}

// Tile rescan -> cumulative offspring counts -> ancestor fill, for one tile.  KIND/PLAIN as
// offspring_of; PLAIN also means no debug weights and no offspring counts.  Thread t owns
// the 4 consecutive particles 4t..4t+3 of the tile, so the tile scan is a thread-serial scan
// plus one block scan.  Every particle's output slots [O_{j-1}, O_j) are written by its own
// thread (neighbouring threads write neighbouring ranges); a thread whose range exceeds
// kDirectSlots hands it to the whole block, which keeps degenerate weights (one particle
// taking every slot) balanced.  For MULTINOMIAL it writes the nondecreasing prefix sums W to
// Wbuf instead.  All threads of the block call it; it returns at once when !in.rs.
template <int MODE, int KIND = -1, bool PLAIN = false>
__device__ __forceinline__ void scan_offspring_tile_in(const RsParams& p, int b,
                                                       const RawW<MODE> raw[kTileItems],
                                                       const TileIn& in) {
  if (!in.rs) return;
  const int kind = KIND >= 0 ? KIND : p.kind;
  double* const w_debug64 = PLAIN ? nullptr : p.w_debug64;
  int32_t* const counts_out = PLAIN ? nullptr : p.counts_out;
  const double* const uvec = PLAIN ? nullptr : p.u;
  __shared__ int olast[kRsThreads];
  __shared__ int heavy[kRsThreads][kTileItems + 2];  // {start, O_0..O_3, owner}
  __shared__ int nheavy;
  const int n = p.n;
  const int nb = (n + kTile - 1) / kTile;
  const int tile0 = b * kTile;
  const int t = threadIdx.x;
  const int cnt = min(kTile, n - tile0);
  double W[kTileItems], P, Pnext, Wn;
  if (t == 0) nheavy = 0;
  tile_cum_weights<MODE>(p, b, raw, in, w_debug64, W, &P, &Pnext, &Wn);
  const uint32_t step = in.step;
  if (kind == SSM_MULTINOMIAL) {
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      const int pos = kTileItems * t + k;
      if (pos < cnt) p.Wbuf[tile0 + pos] = (pos == cnt - 1) ? Pnext : fmin(W[k], Pnext);
    }
    return;
  }
  const double u0 = (kind == SSM_STRATIFIED)
                        ? 0.0
                        : (uvec ? uvec[0] : resample_uniform(rs_seed(p), step, 0u));
  const int o_prev = (b == 0) ? 0 : offspring_of<KIND, PLAIN>(p, P, Wn, u0, step);
  const int o_end = (b == nb - 1) ? n : offspring_of<KIND, PLAIN>(p, Pnext, Wn, u0, step);
  int O[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int pos = kTileItems * t + k;
    if (pos >= cnt - 1)
      O[k] = o_end;  // the tile's last particle closes its range; positions past n own none
    else
      O[k] = min(max(offspring_of<KIND, PLAIN>(p, W[k], Wn, u0, step), o_prev), o_end);
  }
  olast[t] = O[kTileItems - 1];
  __syncthreads();
  const int start = (t == 0) ? o_prev : olast[t - 1];
  if (counts_out) {  // particle k owns the output slots [O[k-1], O[k])
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      const int pos = kTileItems * t + k;
      if (pos < cnt) counts_out[tile0 + pos] = O[k] - (k == 0 ? start : O[k - 1]);
    }
  }
  if (O[kTileItems - 1] - start <= kDirectSlots) {
    int s0 = start;
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      for (int s = s0; s < O[k]; ++s) p.anc[s] = tile0 + kTileItems * t + k;
      s0 = max(s0, O[k]);
    }
  } else {
    const int h = atomicAdd(&nheavy, 1);
    heavy[h][0] = start;
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) heavy[h][k + 1] = O[k];
    heavy[h][kTileItems + 1] = t;
  }
  __syncthreads();
  // heavy ranges: the whole block strides over each one
  const int nh = nheavy;
  for (int h = 0; h < nh; ++h) {
    const int hs = heavy[h][0], he = heavy[h][kTileItems];
    const int owner = heavy[h][kTileItems + 1];
    for (int s = hs + t; s < he; s += kRsThreads) {
      int k = 0;
      while (k < kTileItems - 1 && heavy[h][k + 1] <= s) ++k;
      p.anc[s] = tile0 + kTileItems * owner + k;
    }
  }
}

// The resample of tile b with its inputs from memory: the decision and flags from the
// filter's scalars (or always, for explicit weights), the prefixes and scales from the
// finish kernel.  Every global input is requested before the decision is known (all the
// addresses are in bounds), so the kernel pays one round trip before its arithmetic.
template <int MODE, int KIND = -1, bool PLAIN = false>
__device__ __forceinline__ void scan_offspring_tile(const RsParams& p, int b) {
  const int nb = (p.n + kTile - 1) / kTile;
  RawW<MODE> raw[kTileItems];
  load_tile_raw<MODE>(p, b, raw);
  constexpr bool kLog = (MODE == kWeightsLogF32 || MODE == kWeightsLogF64);
  TileIn in;
  in.P = p.prefix[b];
  in.Pnext = p.prefix[b + 1];
  in.Wn = p.prefix[nb];
  in.m_b = kLog ? p.blk_m[b] : 0.0;
  in.c_b = kLog ? p.blk_c[b] : 0.0;
  in.rs = rs_active(p);
  in.typeless = kLog && p.sc->typeless;
  in.step = rs_step(p);
  scan_offspring_tile_in<MODE, KIND, PLAIN>(p, b, raw, in);
}

// Multinomial: a_k = min{j : W_j >= u_k * W_n} by binary search over Wbuf.
__device__ __forceinline__ void multinomial_search_item(const RsParams& p, int k) {
  if (!rs_active(p)) return;
  if (k >= p.n) return;
  const uint32_t step = rs_step(p);
  const int nb = (p.n + kTile - 1) / kTile;
  const bool typeless = (p.mode == kWeightsLogF32 || p.mode == kWeightsLogF64) && p.sc->typeless;
  const double Wn = typeless ? 1.0 : p.prefix[nb];
  const double uk = p.u ? p.u[k] : resample_uniform(rs_seed(p), step, (uint32_t)k);
  const double q = uk * Wn;
  int lo = 0, hi = p.n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (p.Wbuf[mid] < q)
      lo = mid + 1;
    else
      hi = mid;
  }
  p.anc[k] = min(lo, p.n - 1);
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
  double W[kTileItems], P, Pnext, Wt;
  tile_cum_weights<MODE>(p, b, raw, in, nullptr, W, &P, &Pnext, &Wt);
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
