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
// 4 consecutive particles, a wave scan and one barrier), turns the cumulative weights into
// cumulative offspring counts O_j, and each wave writes its own particles' contiguous range
// of output slots through LDS so that every store instruction covers 64 consecutive slots;
// a particle whose range covers a whole fill window is written directly, so degenerate
// weights stay balanced.
// The FP64 cumulative weights never reach HBM except for Multinomial, whose queries are
// unordered and search them (three levels: tile, group of 16, position).  Neighbouring tiles
// compute their shared boundary count from the same stored prefix, and every cumulative
// sequence is nondecreasing by construction, so the tiles' output ranges partition [0, n)
// exactly.
#include <mutex>

#include "resample_dev.h"

namespace ssm {

constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }

constexpr size_t scratch_bytes_for(size_t n) {
  const size_t nb = (n + kTile - 1) / kTile;
  const size_t ng = (nb + kFinishGroup - 1) / kFinishGroup;
  return round256(nb * kPartial * 8) + round256(nb * 8) +
         round256((nb + 1) * 8) + round256(ng * kGroupRec * 8) + round256(nb * kTile * 8) +
         round256(nb * kMidPerTile * 8) + round256(nb * kRootPerTile * 8);
}

size_t resample_scratch_bytes(int n) { return scratch_bytes_for((size_t)n); }

void resample_bind_scratch(ResampleArgs& a, void* scratch, int n) {
  const size_t nb = (size_t)(n + kTile - 1) / kTile;
  auto r = [](size_t b) { return (b + 255) / 256 * 256; };
  char* p = (char*)scratch;
  a.partials = (double*)p;
  p += r(nb * kPartial * 8);
  a.blk_m = a.partials + 2 * nb;  // the partials' m1 row: the step kernel wrote m1_b there
  a.blk_c = (double*)p;
  p += r(nb * 8);
  a.prefix = (double*)p;
  p += r((nb + 1) * 8);
  a.groups = (double*)p;
  p += r((nb + kFinishGroup - 1) / kFinishGroup * kGroupRec * 8);
  a.Wbuf = (double*)p;
  p += r(nb * kTile * 8);
  a.Wmid = (double*)p;
  p += r(nb * kMidPerTile * 8);
  a.Wroot = (double*)p;
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
    double ex = wave_shfl_up(wave_scan_dpp(cs), 1);
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

// Main kernel: one block per tile (see scan_offspring_tile).  KIND/PLAIN: the particle
// filter's specialisations (a fixed resampler, Philox uniforms keyed from device memory,
// the filter's scalars, weights of whole tiles in a 16-byte aligned buffer, no debug
// outputs).
// The pointers and count of a tile's first loads lead the arguments as scalars, which the
// hardware preloads into registers (the build asks for it).
template <int MODE, int KIND = -1, bool PLAIN = false>
__global__ void __launch_bounds__(kRsThreads)
    k_scan_offspring(const void* w, const BfScalars* sc, const uint64_t* seed_dev,
                     double* prefix, const double* blk_m, const double* blk_c, int n,
                     RsParams p) {
  RsParams q = p;
  q.w = w;
  q.sc = sc;
  q.seed_dev = seed_dev;
  q.prefix = prefix;
  q.blk_m = blk_m;
  q.blk_c = blk_c;
  q.n = n;
  scan_offspring_tile<MODE, KIND, PLAIN>(q, blockIdx.x);
}

// k_scan_offspring's arguments for a parameter block
#define RS_ARGS(rp) rp.w, rp.sc, rp.seed_dev, rp.prefix, rp.blk_m, rp.blk_c, rp.n, rp

template <int MODE>
void launch_scan_offspring(hipStream_t stream, int nb, const RsParams& rp) {
  const dim3 g(nb), blk(kRsThreads);
  // a filter's scalars and key; its weights are Npad (whole tiles) entries
  const bool plain = rp.u == nullptr && rp.w_debug64 == nullptr && rp.counts_out == nullptr &&
                     rp.sc != nullptr && rp.seed_dev != nullptr &&
                     ((uintptr_t)rp.w & 15) == 0;
  if (plain && rp.kind == SSM_SYSTEMATIC)
    hipLaunchKernelGGL((k_scan_offspring<MODE, SSM_SYSTEMATIC, true>), g, blk, 0, stream,
                       RS_ARGS(rp));
  else if (plain && rp.kind == SSM_SYSTEMATIC_MURRAY)
    hipLaunchKernelGGL((k_scan_offspring<MODE, SSM_SYSTEMATIC_MURRAY, true>), g, blk, 0, stream,
                       RS_ARGS(rp));
  else if (plain && rp.kind == SSM_STRATIFIED)
    hipLaunchKernelGGL((k_scan_offspring<MODE, SSM_STRATIFIED, true>), g, blk, 0, stream,
                       RS_ARGS(rp));
  else if (plain && rp.kind == SSM_MULTINOMIAL)
    hipLaunchKernelGGL((k_scan_offspring<MODE, SSM_MULTINOMIAL, true>), g, blk, 0, stream,
                       RS_ARGS(rp));
  else
    hipLaunchKernelGGL((k_scan_offspring<MODE>), g, blk, 0, stream, RS_ARGS(rp));
}

// Multinomial queries: one block per 1024 queries (PLAIN as for k_scan_offspring).
template <bool PLAIN>
__global__ void __launch_bounds__(kRsThreads) k_multinomial_search(RsParams p) {
  multinomial_search_block<PLAIN>(p, blockIdx.x);
}

// The kernels' parameter block for one resample call.
static RsParams rs_params_of(const ResampleArgs& a) {
  RsParams rp;
  rp.kind = a.kind;
  rp.n = a.n;
  rp.mode = a.weight_mode;
  rp.w = a.w;
  rp.u = a.u;
  rp.seed = a.seed;
  rp.seed_dev = a.seed_dev;
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
  rp.Wmid = a.Wmid;
  rp.Wroot = a.Wroot;
  rp.counts_out = (a.kind == SSM_MULTINOMIAL) ? nullptr : a.counts_out;
  return rp;
}

template <int MODE>
__global__ void __launch_bounds__(kRsThreads)
    k_categorical_one(RsParams p, double u, uint32_t stream, int32_t* out) {
  categorical_one_block<MODE>(p, u, stream, out);
}

int launch_categorical_one(ssm_ctx* ctx, hipStream_t stream, const ResampleArgs& a, double u,
                           uint32_t philox_stream, int32_t* out) {
  const RsParams rp = rs_params_of(a);
  if (a.weight_mode == kWeightsLogF32)
    hipLaunchKernelGGL(k_categorical_one<kWeightsLogF32>, dim3(1), dim3(kRsThreads), 0, stream,
                       rp, u, philox_stream, out);
  else if (a.weight_mode == kWeightsLogF64)
    hipLaunchKernelGGL(k_categorical_one<kWeightsLogF64>, dim3(1), dim3(kRsThreads), 0, stream,
                       rp, u, philox_stream, out);
  else
    return set_error(ctx, SSM_ERR_ARG, "categorical draw: log-weight modes only");
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}

int launch_resample(ssm_ctx* ctx, hipStream_t stream, const ResampleArgs& a) {
  if (a.n < 1) return set_error(ctx, SSM_ERR_ARG, "resample: n must be >= 1 (got %d)", a.n);
  if (a.kind < SSM_SYSTEMATIC || a.kind > SSM_MULTINOMIAL)
    return set_error(ctx, SSM_ERR_ARG, "resample: unknown resampler kind %d", a.kind);
  const RsParams rp = rs_params_of(a);
  const int nb = (a.n + kTile - 1) / kTile;
  const dim3 g(nb), blk(kRsThreads);
  switch (a.weight_mode) {
    case kWeightsF64:
      hipLaunchKernelGGL(k_tile_sums<kWeightsF64>, g, blk, 0, stream, rp);
      hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(256), 0, stream, nb, rp.tile_sum,
                         rp.prefix);
      hipLaunchKernelGGL(k_scan_offspring<kWeightsF64>, g, blk, 0, stream, RS_ARGS(rp));
      break;
    case kWeightsF32:
      hipLaunchKernelGGL(k_tile_sums<kWeightsF32>, g, blk, 0, stream, rp);
      hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(256), 0, stream, nb, rp.tile_sum,
                         rp.prefix);
      hipLaunchKernelGGL(k_scan_offspring<kWeightsF32>, g, blk, 0, stream, RS_ARGS(rp));
      break;
    case kWeightsLogF32:
      launch_scan_offspring<kWeightsLogF32>(stream, nb, rp);
      break;
    case kWeightsLogF64:
      launch_scan_offspring<kWeightsLogF64>(stream, nb, rp);
      break;
    default: return set_error(ctx, SSM_ERR_ARG, "resample: bad weight mode");
  }
  if (a.kind == SSM_MULTINOMIAL) {
    if (rp.u == nullptr && rp.sc != nullptr && rp.seed_dev != nullptr)
      hipLaunchKernelGGL(k_multinomial_search<true>, g, blk, 0, stream, rp);
    else
      hipLaunchKernelGGL(k_multinomial_search<false>, g, blk, 0, stream, rp);
  }
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

// ---- scratch of the entry point that allocates nothing ----------------------------------
// ssm_sample_ancestors_noalloc's scratch is this file-scope device array: one copy per GPU,
// held until the process ends, shared by every call on that GPU.  Host code never passes
// the array's name (a host-side stand-in) to a kernel: each device's address comes from
// hipGetSymbolAddress, taken once per device when a context is created on it.
constexpr int kNoAllocMaxN = 1 << 20;
constexpr size_t kNoAllocScratchBytes = scratch_bytes_for(kNoAllocMaxN);
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kMaxDevices] = {};

int noalloc_setup(ssm_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_error(ctx, SSM_ERR_ARG, "device %d beyond the %d supported", ctx->device,
                     kMaxDevices);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[ctx->device]) return SSM_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));  // device is current
  if (e != hipSuccess || !ptr)
    return set_error(ctx, SSM_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                     hipGetErrorString(e));
  g_noalloc_ptr[ctx->device] = ptr;
  return SSM_OK;
}

}  // namespace ssm

using namespace ssm;

extern "C" int32_t ssm_sample_ancestors_noalloc(ssm_ctx* ctx, int32_t kind, int32_t n,
                                                const void* d_w, int32_t w_dtype,
                                                const double* d_u, uint64_t seed,
                                                uint32_t step, int32_t* d_ancestors) {
  if (!ctx) return SSM_ERR_ARG;
  if (n < 1 || !d_w || !d_ancestors)
    return set_error(ctx, SSM_ERR_ARG, "ssm_sample_ancestors_noalloc: bad arguments (n=%d)", n);
  if (n > kNoAllocMaxN)
    return set_error(ctx, SSM_ERR_ARG,
                     "ssm_sample_ancestors_noalloc: n=%d exceeds its maximum of %d weights "
                     "(use ssm_sample_ancestors)", n, kNoAllocMaxN);
  if (w_dtype != SSM_F32 && w_dtype != SSM_F64)
    return set_error(ctx, SSM_ERR_ARG, "ssm_sample_ancestors_noalloc: bad w_dtype %d", w_dtype);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  void* ptr;
  {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    ptr = g_noalloc_ptr[ctx->device];
  }
  if (!ptr)
    return set_error(ctx, SSM_ERR_STATE, "ssm_sample_ancestors_noalloc: no scratch address");
  // every scratch word the kernels read is written earlier in the same call (tile sums,
  // prefixes, cumulative weights), so nothing needs clearing between calls
  ResampleArgs a;
  a.kind = kind;
  a.n = n;
  a.weight_mode = (w_dtype == SSM_F64) ? kWeightsF64 : kWeightsF32;
  a.w = d_w;
  a.u = d_u;
  a.seed = seed;
  a.step_host = step;
  a.anc = d_ancestors;
  resample_bind_scratch(a, ptr, n);
  return launch_resample(ctx, ctx->stream, a);
}

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
