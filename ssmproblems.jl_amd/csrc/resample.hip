// resample.hip -- ancestor sampling on gfx950: Systematic (CPU and GPU definitions of the
// reference), Stratified and Multinomial, plus the offspring<->ancestor conversions.
//
// Reference semantics (paths relative to the reference checkout):
//   Systematic (CPU)  GF/src/resamplers.jl:216-238  a_k = min{j : n*W_j > u0 + (k-1)}
//   Systematic (GPU)  GF/ext/CUDAExt.jl:49-58       O_j = min(n, floor(n*W_j/W_n + u0))
//   Stratified        GF/ext/CUDAExt.jl:68-78       O_j = min(n, floor(r_j + u[min(n, floor(r_j+1))]))
//   Multinomial       GF/ext/CUDAExt.jl:31-39       a_k = min{j : W_j >= u_k W_n}
//   offspring->ancestors CUDAExt.jl:90-117, ancestors->offspring CUDAExt.jl:119-142
// with W the inclusive prefix sum of the weights, taken here in FP64.
//
// Design.  Weights are cut into tiles of 1024 (see pf_common.h).  Tile prefixes come either
// from a particle filter's finish kernel (its tile partials already hold each tile's
// weight sum) or, for explicit weight vectors, from a tile-sum kernel plus a one-block
// scan.  The main kernel rescans one tile in FP64 on top of its prefix (thread-serial over
// 4 consecutive particles, then one block scan), turns the prefix sums into cumulative
// offspring counts O_j, and writes each particle's output slots [O_{j-1}, O_j) from the
// thread that owns it; ranges longer than 32 slots are written by the whole block, so
// degenerate weights stay balanced.
// The FP64 prefix sums never reach HBM except for Multinomial, whose queries are unordered
// and need a global binary search.  Neighbouring tiles compute their shared boundary count
// from the same stored prefix, and every prefix sequence is made nondecreasing with an exact
// running max, so the tiles' output ranges partition [0, n) exactly.
#include "device_common.h"
#include "internal.h"
#include "pf_common.h"

namespace ssm {

constexpr int kRsThreads = kTileThreads;

size_t resample_scratch_bytes(int n) {
  const size_t nb = (size_t)(n + kTile - 1) / kTile;
  auto r = [](size_t b) { return (b + 255) / 256 * 256; };
  return r(nb * kPartial * 8) + r(nb * 8) + r(nb * 8) + r((nb + 1) * 8) + r((size_t)n * 8);
}

void resample_bind_scratch(ResampleArgs& a, void* scratch, int n) {
  const size_t nb = (size_t)(n + kTile - 1) / kTile;
  auto r = [](size_t b) { return (b + 255) / 256 * 256; };
  char* p = (char*)scratch;
  a.partials = (double*)p;
  p += r(nb * kPartial * 8);
  a.blk_m = (double*)p;
  p += r(nb * 8);
  a.blk_c = (double*)p;
  p += r(nb * 8);
  a.prefix = (double*)p;
  p += r((nb + 1) * 8);
  a.Wbuf = (double*)p;
}

struct RsParams {
  int kind, n, mode;
  const void* w;
  const double* u;
  uint64_t seed;
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
};

__device__ __forceinline__ bool rs_active(const RsParams& p) {
  return p.sc == nullptr || p.threshold_n >= p.sc->ess;
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

// Inclusive FP64 scan across the block (thread order), fixed order; *tot = block total.
__device__ __forceinline__ double block_incl_scan_d(double v, double* lds, double* tot) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const double incl = wave_inclusive_scan(v);
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
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double o = wave_shfl_up(v, d);
    if (lane >= d) v = fmax(v, o);
  }
  __syncthreads();
  if (lane == kWave - 1) lds[w] = v;
  __syncthreads();
  for (int k = 0; k < w; ++k) v = fmax(v, lds[k]);
  *tot = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
  return v;
}

// Explicit weights, kernel 1: FP64 tile sums (thread-serial over its striped items, then
// the block in fixed order).
template <int MODE>
__global__ void __launch_bounds__(kRsThreads) k_tile_sums(RsParams p) {
  __shared__ double lds[kRsThreads / kWave];
  const int b = blockIdx.x, tile0 = b * kTile;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int j = tile0 + k * kRsThreads + threadIdx.x;
    if (j < p.n) s += rs_weight<MODE>(p, j, b);
  }
  double tot;
  (void)block_incl_scan_d(s, lds, &tot);
  if (threadIdx.x == 0) p.tile_sum[b] = tot;
}

// Explicit weights, kernel 2 (one block): exclusive scan of the tile sums into
// nondecreasing prefixes, prefix[nb] = total.
__global__ void __launch_bounds__(256) k_tile_prefix(int nb, const double* tile_sum,
                                                     double* prefix) {
  __shared__ double lds[4];
  const int tid = threadIdx.x;
  const int chunk = (nb + 255) / 256;
  const int lo = min(nb, tid * chunk), hi = min(nb, lo + chunk);
  double cs = 0.0;
  for (int k = lo; k < hi; ++k) cs += tile_sum[k];
  double tot;
  (void)block_incl_scan_d(cs, lds, &tot);  // leaves the wave totals in lds
  // exclusive start of this thread's chunk, from the same wave scan and wave totals
  double run;
  {
    const int w = tid / kWave, lane = tid & (kWave - 1);
    double ex = wave_shfl_up(wave_inclusive_scan(cs), 1);
    if (lane == 0) ex = 0.0;
    for (int k = 0; k < w; ++k) ex += lds[k];
    run = ex;
  }
  double pm = -INFINITY;
  for (int k = lo; k < hi; ++k) {
    pm = fmax(pm, run);
    prefix[k] = pm;
    run += tile_sum[k];
  }
  double mx;
  __syncthreads();
  const double carry_incl = block_incl_max_d(pm, lds, &mx);
  double carry = wave_shfl_up(carry_incl, 1);
  if ((tid & 63) == 0) carry = -INFINITY;
  for (int k = 0; k < tid / kWave; ++k) carry = fmax(carry, lds[k]);
  for (int k = lo; k < hi; ++k) prefix[k] = fmax(prefix[k], carry);
  if (tid == 0) prefix[nb] = fmax(tot, mx);
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

__device__ __forceinline__ int offspring_of(const RsParams& p, double W, double Wn, double u0,
                                            uint32_t step) {
  const int n = p.n;
  if (p.kind == SSM_SYSTEMATIC) return offspring_systematic(W, u0, n);
  const double r = ((double)n * W) / Wn;
  if (p.kind == SSM_SYSTEMATIC_MURRAY) return clamp_floor(r + u0, n);
  // STRATIFIED: k = min(n, floor(r + 1)) is 1-based into u.
  int k = clamp_floor(r + 1.0, n);
  if (k < 1) k = 1;
  const double uk = p.u ? p.u[k - 1] : resample_uniform(p.seed, step, (uint32_t)(k - 1));
  return clamp_floor(r + uk, n);
}

// Exclusive FP64 scan of one value per thread across the block (thread order, fixed
// order); *tot = block total.
__device__ __forceinline__ double block_excl_scan_d(double v, double* lds, double* tot) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const double incl = wave_inclusive_scan(v);
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
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double o = wave_shfl_up(v, d);
    if (lane >= d) v = fmax(v, o);
  }
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
template <int MODE>
__global__ void __launch_bounds__(kRsThreads) k_scan_offspring(RsParams p) {
  if (!rs_active(p)) return;
  __shared__ double lds[kRsThreads / kWave];
  __shared__ int olast[kRsThreads];
  __shared__ int heavy[kRsThreads][kTileItems + 2];  // {start, O_0..O_3, owner}
  __shared__ int nheavy;
  const int n = p.n;
  const int b = blockIdx.x;
  const int nb = (n + kTile - 1) / kTile;
  const int tile0 = b * kTile;
  const int cnt = min(kTile, n - tile0);
  const int t = threadIdx.x;
  const bool typeless = (MODE == kWeightsLogF32 || MODE == kWeightsLogF64) && p.sc->typeless;
  double P, Pnext, Wn;
  if (typeless) {
    P = (double)tile0 / (double)n;
    Pnext = (b == nb - 1) ? 1.0 : (double)(tile0 + kTile) / (double)n;
    Wn = 1.0;
  } else {
    P = p.prefix[b];
    Pnext = p.prefix[b + 1];
    Wn = p.prefix[nb];
  }
  if (t == 0) nheavy = 0;
  // thread-serial scan of 4 consecutive weights, then one block scan of the thread sums
  double W[kTileItems];
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int j = tile0 + kTileItems * t + k;
    double wj = 0.0;
    if (j < n) {
      wj = rs_weight<MODE>(p, j, b);
      if (p.w_debug64) p.w_debug64[j] = wj;
    }
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
  const uint32_t step = rs_step(p);
  if (p.kind == SSM_MULTINOMIAL) {
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      const int pos = kTileItems * t + k;
      if (pos < cnt) p.Wbuf[tile0 + pos] = (pos == cnt - 1) ? Pnext : fmin(W[k], Pnext);
    }
    return;
  }
  const double u0 = (p.kind == SSM_STRATIFIED)
                        ? 0.0
                        : (p.u ? p.u[0] : resample_uniform(p.seed, step, 0u));
  const int o_prev = (b == 0) ? 0 : offspring_of(p, P, Wn, u0, step);
  const int o_end = (b == nb - 1) ? n : offspring_of(p, Pnext, Wn, u0, step);
  int O[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int pos = kTileItems * t + k;
    if (pos >= cnt - 1)
      O[k] = o_end;  // the tile's last particle closes its range; positions past n own none
    else
      O[k] = min(max(offspring_of(p, W[k], Wn, u0, step), o_prev), o_end);
  }
  olast[t] = O[kTileItems - 1];
  __syncthreads();
  const int start = (t == 0) ? o_prev : olast[t - 1];
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

// Multinomial: a_k = min{j : W_j >= u_k * W_n} by binary search over Wbuf.
__global__ void __launch_bounds__(kRsThreads) k_multinomial_search(RsParams p) {
  if (!rs_active(p)) return;
  const int k = blockIdx.x * kRsThreads + threadIdx.x;
  if (k >= p.n) return;
  const uint32_t step = rs_step(p);
  const int nb = (p.n + kTile - 1) / kTile;
  const bool typeless = (p.mode == kWeightsLogF32 || p.mode == kWeightsLogF64) && p.sc->typeless;
  const double Wn = typeless ? 1.0 : p.prefix[nb];
  const double uk = p.u ? p.u[k] : resample_uniform(p.seed, step, (uint32_t)k);
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

int launch_resample(ssm_ctx* ctx, hipStream_t stream, const ResampleArgs& a) {
  if (a.n < 1) return set_error(ctx, SSM_ERR_ARG, "resample: n must be >= 1 (got %d)", a.n);
  if (a.kind < SSM_SYSTEMATIC || a.kind > SSM_MULTINOMIAL)
    return set_error(ctx, SSM_ERR_ARG, "resample: unknown resampler kind %d", a.kind);
  RsParams rp;
  rp.kind = a.kind;
  rp.n = a.n;
  rp.mode = a.weight_mode;
  rp.w = a.w;
  rp.u = a.u;
  rp.seed = a.seed;
  rp.step_host = a.step_host;
  rp.sc = a.sc;
  rp.threshold_n = a.threshold_n;
  rp.anc = a.anc;
  rp.w_debug64 = a.w_debug64;
  rp.blk_m = a.blk_m;
  rp.blk_c = a.blk_c;
  rp.tile_sum = a.partials;
  rp.prefix = a.prefix;
  rp.Wbuf = a.Wbuf;
  const int nb = (a.n + kTile - 1) / kTile;
  const dim3 g(nb), blk(kRsThreads);
  switch (a.weight_mode) {
    case kWeightsF64:
      hipLaunchKernelGGL(k_tile_sums<kWeightsF64>, g, blk, 0, stream, rp);
      hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(256), 0, stream, nb, rp.tile_sum,
                         rp.prefix);
      hipLaunchKernelGGL(k_scan_offspring<kWeightsF64>, g, blk, 0, stream, rp);
      break;
    case kWeightsF32:
      hipLaunchKernelGGL(k_tile_sums<kWeightsF32>, g, blk, 0, stream, rp);
      hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(256), 0, stream, nb, rp.tile_sum,
                         rp.prefix);
      hipLaunchKernelGGL(k_scan_offspring<kWeightsF32>, g, blk, 0, stream, rp);
      break;
    case kWeightsLogF32:
      hipLaunchKernelGGL(k_scan_offspring<kWeightsLogF32>, g, blk, 0, stream, rp);
      break;
    case kWeightsLogF64:
      hipLaunchKernelGGL(k_scan_offspring<kWeightsLogF64>, g, blk, 0, stream, rp);
      break;
    default: return set_error(ctx, SSM_ERR_ARG, "resample: bad weight mode");
  }
  if (a.kind == SSM_MULTINOMIAL)
    hipLaunchKernelGGL(k_multinomial_search, dim3((a.n + kRsThreads - 1) / kRsThreads),
                       dim3(kRsThreads), 0, stream, rp);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}

// ---- offspring <-> ancestors ------------------------------------------------------
// One thread per output slot s: ancestor = first i with O_i > s (O nondecreasing), which
// equals the reference's range scatter for valid input and needs no load balancing.
__global__ void __launch_bounds__(kRsThreads)
    k_offspring_to_ancestors(int n, const int32_t* __restrict__ O, int32_t* __restrict__ anc) {
  const int s = blockIdx.x * kRsThreads + threadIdx.x;
  if (s >= n) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (O[mid] > s)
      hi = mid;
    else
      lo = mid + 1;
  }
  anc[s] = min(lo, n - 1);
}

__global__ void __launch_bounds__(kRsThreads) k_fill_i32(int n, int32_t* p, int32_t v) {
  const int i = blockIdx.x * kRsThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kRsThreads)
    k_ancestors_to_offspring(int n, const int32_t* __restrict__ anc, int32_t* off) {
  const int i = blockIdx.x * kRsThreads + threadIdx.x;
  if (i >= n) return;
  const int a = anc[i];
  if (a >= 0 && a < n) atomicAdd(&off[a], 1);
}

}  // namespace ssm

using namespace ssm;

extern "C" int32_t ssm_sample_ancestors(ssm_ctx* ctx, int32_t kind, int32_t n, const void* d_w,
                                        int32_t w_dtype, const double* d_u, uint64_t seed,
                                        uint32_t step, int32_t* d_ancestors) {
  if (!ctx) return SSM_ERR_ARG;
  if (n < 1 || !d_w || !d_ancestors)
    return set_error(ctx, SSM_ERR_ARG, "ssm_sample_ancestors: bad arguments (n=%d)", n);
  if (w_dtype != SSM_F32 && w_dtype != SSM_F64)
    return set_error(ctx, SSM_ERR_ARG, "ssm_sample_ancestors: bad w_dtype %d", w_dtype);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const size_t need = resample_scratch_bytes(n);
  if (need > ctx->rs_scratch_bytes) {
    SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->rs_scratch) (void)hipFree(ctx->rs_scratch);
    ctx->rs_scratch = nullptr;
    ctx->rs_scratch_bytes = 0;
    if (hipMalloc(&ctx->rs_scratch, need) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "ssm_sample_ancestors: cannot allocate %zu B", need);
    ctx->rs_scratch_bytes = need;
  }
  ResampleArgs a;
  a.kind = kind;
  a.n = n;
  a.weight_mode = (w_dtype == SSM_F64) ? kWeightsF64 : kWeightsF32;
  a.w = d_w;
  a.u = d_u;
  a.seed = seed;
  a.step_host = step;
  a.anc = d_ancestors;
  resample_bind_scratch(a, ctx->rs_scratch, n);
  return launch_resample(ctx, ctx->stream, a);
}

extern "C" int32_t ssm_sample_ancestors_host(ssm_ctx* ctx, int32_t kind, int32_t n,
                                             const double* w, const double* u,
                                             int32_t* ancestors) {
  if (!ctx) return SSM_ERR_ARG;
  if (n < 1 || !w || !u || !ancestors)
    return set_error(ctx, SSM_ERR_ARG, "ssm_sample_ancestors_host: bad arguments (n=%d)", n);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if ((size_t)n > ctx->rs_host_cap) {
    SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(ctx->rs_w);
    (void)hipFree(ctx->rs_u);
    (void)hipFree(ctx->rs_a);
    ctx->rs_w = nullptr;
    ctx->rs_u = nullptr;
    ctx->rs_a = nullptr;
    ctx->rs_host_cap = 0;
    if (hipMalloc(&ctx->rs_w, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&ctx->rs_u, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&ctx->rs_a, sizeof(int32_t) * n) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "ssm_sample_ancestors_host: allocation failed");
    ctx->rs_host_cap = n;
  }
  const size_t nu = (kind == SSM_SYSTEMATIC || kind == SSM_SYSTEMATIC_MURRAY) ? 1 : (size_t)n;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->rs_w, w, sizeof(double) * n, hipMemcpyHostToDevice,
                                    ctx->stream));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->rs_u, u, sizeof(double) * nu, hipMemcpyHostToDevice,
                                    ctx->stream));
  int rc = ssm_sample_ancestors(ctx, kind, n, ctx->rs_w, SSM_F64, ctx->rs_u, 0, 0, ctx->rs_a);
  if (rc) return rc;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(ancestors, ctx->rs_a, sizeof(int32_t) * n,
                                    hipMemcpyDeviceToHost, ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return SSM_OK;
}

extern "C" int32_t ssm_offspring_to_ancestors(ssm_ctx* ctx, int32_t n,
                                              const int32_t* d_offspring,
                                              int32_t* d_ancestors) {
  if (!ctx) return SSM_ERR_ARG;
  if (n < 1 || !d_offspring || !d_ancestors)
    return set_error(ctx, SSM_ERR_ARG, "ssm_offspring_to_ancestors: bad arguments");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_offspring_to_ancestors, dim3((n + kRsThreads - 1) / kRsThreads),
                     dim3(kRsThreads), 0, ctx->stream, n, d_offspring, d_ancestors);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}

extern "C" int32_t ssm_ancestors_to_offspring(ssm_ctx* ctx, int32_t n,
                                              const int32_t* d_ancestors,
                                              int32_t* d_offspring) {
  if (!ctx) return SSM_ERR_ARG;
  if (n < 1 || !d_offspring || !d_ancestors)
    return set_error(ctx, SSM_ERR_ARG, "ssm_ancestors_to_offspring: bad arguments");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const dim3 g((n + kRsThreads - 1) / kRsThreads);
  hipLaunchKernelGGL(k_fill_i32, g, dim3(kRsThreads), 0, ctx->stream, n, d_offspring, 0);
  hipLaunchKernelGGL(k_ancestors_to_offspring, g, dim3(kRsThreads), 0, ctx->stream, n,
                     d_ancestors, d_offspring);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}
