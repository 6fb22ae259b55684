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
// Design: tiles of kTile = 2048 weights (256 threads x 8).  Kernel 1 reduces each tile in
// FP64 and the last block to finish scans the tile sums (fixed order: deterministic).
// Kernel 2 rescans each tile with its prefix, turns prefix sums into cumulative offspring
// counts O_j, keeps them in LDS and fills the tile's contiguous output range
// [O_{tile start}, O_{tile end}) with one thread per output slot (binary search in LDS),
// so the ancestor write is balanced and coalesced however uneven the weights are.  The
// FP64 prefix sums never reach HBM except for Multinomial, whose queries are unordered and
// need a global binary search.
#include "device_common.h"
#include "internal.h"

namespace ssm {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;

size_t resample_scratch_bytes(int n) {
  const size_t nb = (size_t)(n + kTile - 1) / kTile;
  size_t b = 0;
  b += 256;                         // counter + total
  b += ((nb * 8 + 255) / 256) * 256;        // tile_sum
  b += (((nb + 1) * 8 + 255) / 256) * 256;  // prefix
  b += (((size_t)n * 8 + 255) / 256) * 256;  // W (multinomial)
  return b;
}

void resample_bind_scratch(ResampleArgs& a, void* scratch, int n) {
  const size_t nb = (size_t)(n + kTile - 1) / kTile;
  char* p = (char*)scratch;
  a.counter = (uint32_t*)p;
  a.total = (double*)(p + 8);
  p += 256;
  a.tile_sum = (double*)p;
  p += ((nb * 8 + 255) / 256) * 256;
  a.prefix = (double*)p;
  p += (((nb + 1) * 8 + 255) / 256) * 256;
  a.Wbuf = (double*)p;
}

// Kernel parameters passed by value.
struct RsParams {
  int kind;
  int n;
  int mode;
  const void* w;
  const double* u;
  uint64_t seed;
  uint32_t step_host;
  BfScalars* sc;
  double threshold_n;
  int32_t* anc;
  float* w_debug;
  double* w_debug64;
  double* tile_sum;
  double* prefix;
  uint32_t* counter;
  double* Wbuf;
  double* total;
};

__device__ __forceinline__ bool rs_active(const RsParams& p) {
  return p.sc == nullptr || p.threshold_n >= p.sc->ess;
}

__device__ __forceinline__ uint32_t rs_step(const RsParams& p) {
  return p.sc ? (uint32_t)p.sc->step : p.step_host;
}

// Weight j as the resampler consumes it (FP64).  For particle-filter log-weights this is
// softmax of the normalised log-weights: exp(nl_j - max nl) / sum_k exp(nl_k - max nl)
// with the normalisers kept in BfScalars (containers.jl:140-141), or 1/n for the lazy
// TypelessZero weights (containers.jl:62-65).
template <int MODE>
__device__ __forceinline__ double rs_weight(const RsParams& p, int j, bool write_dbg) {
  if (MODE == kWeightsF64) return ((const double*)p.w)[j];
  if (MODE == kWeightsF32) return (double)((const float*)p.w)[j];
  const BfScalars* sc = p.sc;
  if (MODE == kWeightsLogF32) {
    float wf;
    if (sc->typeless) {
      const double w = 1.0 / (double)p.n;
      if (write_dbg && p.w_debug64) p.w_debug64[j] = w;
      return w;
    }
    const float nl = ((const float*)p.w)[j] - (float)sc->lse;
    wf = expf(nl - (float)sc->mxn) / (float)sc->s1;
    if (write_dbg && p.w_debug64) p.w_debug64[j] = (double)wf;
    return (double)wf;
  } else {
    double w;
    if (sc->typeless) {
      w = 1.0 / (double)p.n;
    } else {
      const double nl = ((const double*)p.w)[j] - sc->lse;
      w = exp(nl - sc->mxn) / sc->s1;
    }
    if (write_dbg && p.w_debug64) p.w_debug64[j] = w;
    return w;
  }
}

// Block-wide FP64 sum (fixed order), result valid in thread 0.
__device__ __forceinline__ double block_sum_d(double v, double* lds4) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (lane == 0) lds4[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((lds4[0] + lds4[1]) + (lds4[2] + lds4[3]));
  return r;
}

// Block exclusive scan of one FP64 value per thread; returns the exclusive prefix and
// writes the block total to *tot.  Fixed order.
__device__ __forceinline__ double block_exclusive_scan_d(double v, double* lds, double* tot) {
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const double incl = wave_inclusive_scan(v);
  double excl = wave_shfl_up(incl, 1);
  if (lane == 0) excl = 0.0;
  if (lane == kWave - 1) lds[w] = incl;
  __syncthreads();
  double wp = 0.0;
  for (int k = 0; k < w; ++k) wp += lds[k];
  *tot = ((lds[0] + lds[1]) + lds[2]) + lds[3];
  __syncthreads();
  return wp + excl;
}

// Kernel 1: FP64 tile sums; the last block scans them into prefix[0..nb] and total.
template <int MODE>
__global__ void __launch_bounds__(kThreads) k_tile_sums(RsParams p) {
  if (!rs_active(p)) return;
  __shared__ double lds[kThreads / kWave];
  __shared__ int is_last;
  const int nb = (p.n + kTile - 1) / kTile;
  const int b = blockIdx.x;
  const int base = b * kTile + threadIdx.x * kItems;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int j = base + k;
    if (j < p.n) s += rs_weight<MODE>(p, j, true);
  }
  const double bs = block_sum_d(s, lds);
  uint32_t* counter = p.sc ? &p.sc->cnt_scan : p.counter;
  if (threadIdx.x == 0) {
    p.tile_sum[b] = bs;
    __threadfence();
    is_last = (atomicAdd(counter, 1u) == (uint32_t)(nb - 1));
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  // Exclusive scan of the nb tile sums: each thread owns a contiguous chunk.
  const int chunk = (nb + kThreads - 1) / kThreads;
  const int lo = min(nb, (int)threadIdx.x * chunk), hi = min(nb, lo + chunk);
  double cs = 0.0;
  for (int k = lo; k < hi; ++k) cs += __builtin_nontemporal_load(&p.tile_sum[k]);
  __shared__ double lds2[kThreads / kWave];
  double tot;
  double run = block_exclusive_scan_d(cs, lds2, &tot);
  for (int k = lo; k < hi; ++k) {
    p.prefix[k] = run;
    run += __builtin_nontemporal_load(&p.tile_sum[k]);
  }
  if (threadIdx.x == 0) {
    p.prefix[nb] = tot;
    *p.total = tot;
    if (p.sc) p.sc->total_w = tot;
    *counter = 0u;
  }
}

// Cumulative offspring count for an inclusive prefix W (GPU forms use scan total Wn).
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

// Kernel 2: tile rescan -> cumulative offspring in LDS -> balanced ancestor fill.
// For MULTINOMIAL it only writes the prefix sums W to Wbuf.
template <int MODE>
__global__ void __launch_bounds__(kThreads) k_scan_offspring(RsParams p) {
  if (!rs_active(p)) return;
  __shared__ double lds[kThreads / kWave];
  __shared__ int O[kTile];
  const int n = p.n;
  const int b = blockIdx.x;
  const int nb = (n + kTile - 1) / kTile;
  const int tile0 = b * kTile;
  const int cnt = min(kTile, n - tile0);
  const int base = tile0 + threadIdx.x * kItems;
  double w[kItems];
  double ts = 0.0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int j = base + k;
    w[k] = (j < n) ? rs_weight<MODE>(p, j, false) : 0.0;
    ts += w[k];
  }
  double tot;
  const double excl = block_exclusive_scan_d(ts, lds, &tot);
  (void)tot;
  const double P = p.prefix[b];
  const double Pnext = p.prefix[b + 1];
  const double Wn = *p.total;
  const uint32_t step = rs_step(p);
  double acc = P + excl;
  if (p.kind == SSM_MULTINOMIAL) {
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      const int j = base + k;
      acc += w[k];
      if (j < n) p.Wbuf[j] = (j == tile0 + cnt - 1) ? Pnext : acc;
    }
    return;
  }
  const double u0 = (p.kind == SSM_STRATIFIED)
                        ? 0.0
                        : (p.u ? p.u[0] : resample_uniform(p.seed, step, 0u));
  // Tile boundary counts are computed from the shared prefixes so neighbouring tiles
  // agree exactly on where one's output range ends and the next one's begins.
  const int o_prev = (b == 0) ? 0 : offspring_of(p, P, Wn, u0, step);
  const int o_end = (b == nb - 1) ? n : offspring_of(p, Pnext, Wn, u0, step);
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int j = base + k;
    acc += w[k];
    if (j < n) {
      int o;
      if (j == tile0 + cnt - 1)
        o = o_end;
      else
        o = min(max(offspring_of(p, acc, Wn, u0, step), o_prev), o_end);
      O[j - tile0] = o;
    }
  }
  __syncthreads();
  // Slot s belongs to the first local j with O[j] > s.
  for (int s = o_prev + (int)threadIdx.x; s < o_end; s += kThreads) {
    int lo = 0, hi = cnt - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (O[mid] > s)
        hi = mid;
      else
        lo = mid + 1;
    }
    p.anc[s] = tile0 + lo;
  }
}

// Kernel 3 (multinomial): a_k = min{j : W_j >= u_k * W_n} by binary search over Wbuf.
__global__ void __launch_bounds__(kThreads) k_multinomial_search(RsParams p) {
  if (!rs_active(p)) return;
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= p.n) return;
  const uint32_t step = rs_step(p);
  const double Wn = *p.total;
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

template <int MODE>
static void launch_mode(hipStream_t s, const RsParams& rp, int nb) {
  hipLaunchKernelGGL(k_tile_sums<MODE>, dim3(nb), dim3(kThreads), 0, s, rp);
  hipLaunchKernelGGL(k_scan_offspring<MODE>, dim3(nb), dim3(kThreads), 0, s, rp);
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
  rp.w_debug = a.w_debug;
  rp.w_debug64 = a.w_debug64;
  rp.tile_sum = a.tile_sum;
  rp.prefix = a.prefix;
  rp.counter = a.counter;
  rp.Wbuf = a.Wbuf;
  rp.total = a.total;
  const int nb = (a.n + kTile - 1) / kTile;
  switch (a.weight_mode) {
    case kWeightsF64: launch_mode<kWeightsF64>(stream, rp, nb); break;
    case kWeightsF32: launch_mode<kWeightsF32>(stream, rp, nb); break;
    case kWeightsLogF32: launch_mode<kWeightsLogF32>(stream, rp, nb); break;
    case kWeightsLogF64: launch_mode<kWeightsLogF64>(stream, rp, nb); break;
    default: return set_error(ctx, SSM_ERR_ARG, "resample: bad weight mode");
  }
  if (a.kind == SSM_MULTINOMIAL)
    hipLaunchKernelGGL(k_multinomial_search, dim3((a.n + kThreads - 1) / kThreads),
                       dim3(kThreads), 0, stream, rp);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}

// ---- offspring <-> ancestors ------------------------------------------------------
// One thread per output slot s: ancestor = first i with O_i > s (O nondecreasing), which
// equals the reference's range scatter for valid input and needs no load balancing.
__global__ void __launch_bounds__(kThreads)
    k_offspring_to_ancestors(int n, const int32_t* __restrict__ O, int32_t* __restrict__ anc) {
  const int s = blockIdx.x * kThreads + threadIdx.x;
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

__global__ void __launch_bounds__(kThreads) k_fill_i32(int n, int32_t* p, int32_t v) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kThreads)
    k_ancestors_to_offspring(int n, const int32_t* __restrict__ anc, int32_t* off) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
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
    if (ctx->rs_scratch) (void)hipFree(ctx->rs_scratch);
    ctx->rs_scratch = nullptr;
    ctx->rs_scratch_bytes = 0;
    if (hipMalloc(&ctx->rs_scratch, need) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "ssm_sample_ancestors: cannot allocate %zu B", need);
    ctx->rs_scratch_bytes = need;
    // The counter must start at zero; the kernels reset it after each use.
    SSM_HIP_CHECK(ctx, hipMemset(ctx->rs_scratch, 0, 256));
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
  hipLaunchKernelGGL(k_offspring_to_ancestors, dim3((n + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, ctx->stream, n, d_offspring, d_ancestors);
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
  const dim3 g((n + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_fill_i32, g, dim3(kThreads), 0, ctx->stream, n, d_offspring, 0);
  hipLaunchKernelGGL(k_ancestors_to_offspring, g, dim3(kThreads), 0, ctx->stream, n,
                     d_ancestors, d_offspring);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}
