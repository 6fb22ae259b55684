// tree.hip -- sparse path storage of a bootstrap filter's particles on the device.
//
// The device form of ParallelParticleTree (GF/ext/CUDAExt.jl:146-298), with the path
// semantics of ParticleTree (GF/src/containers.jl:466-632) and its callback
// (CUDAExt.jl:279-298): Jacob, Murray & Rubenthaler (2015), "Path storage in the particle
// filter".  Every node holds one particle state, its parent node and its live offspring
// count.  An insert of a new generation
//   1. takes each new particle's parent node as the leaf of its ancestor,
//   2. stores at every current leaf its number of children,
//   3. prunes: a childless leaf is free, and walking up, every node whose count drops to
//      zero is free too,
//   4. assigns the new particles to free slots in index order (the i-th new particle takes
//      the i-th free slot: an exclusive scan of the free flags), doubling the pool first
//      when fewer than N slots are free,
//   5. writes the new states, parents and leaves.
// The node pool keeps states as SoA planes states[d][M] in the filter's type.
//
// Launches per insert: one link kernel (steps 1-3), the free-slot count per scan tile, the
// one-block scan of those counts, and the assign kernel (steps 4-5).  After a systematic
// or stratified resample in ssm_bf_step the children counts come from the resampler, which
// writes them for an attached tree, so step 2 needs no atomics; otherwise (multinomial, or
// steps taken by other entry points) they are counted with atomics in one extra kernel.
#include <algorithm>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "internal.h"

namespace ssm {

constexpr int kTreeThreads = 256;
constexpr int kTreeScanMax = 4096;  // scan tiles per insert, at most (one-block scan)
constexpr int kTreeScanPer = kTreeScanMax / kTreeThreads;

template <typename T>
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_roots(int N, int64_t M, int D, const T* __restrict__ x, int Npad, T* states,
                 int32_t* parents, int32_t* offspring, int32_t* leaves) {
  const int64_t k = (int64_t)blockIdx.x * kTreeThreads + threadIdx.x;
  if (k >= M) return;
  parents[k] = -1;
  offspring[k] = 0;
  if (k < N) {
    for (int d = 0; d < D; ++d) states[(size_t)d * M + k] = x[(size_t)d * Npad + k];
    leaves[k] = (int32_t)k;
  }
}

// Walks up from node p, one of whose children was just freed.  A node whose count falls to
// zero is freed and its parent loses a child in turn.  Counts only fall during a prune, so
// a plain read can be stale but never below the true count; p's count includes the freed
// child until this walk takes it off, so reading 1 means that child was p's last and no
// other walk will reach p: the count is cleared with a store, no atomic round trip.  The
// parent link is read alongside (links do not change during a prune).
__device__ __forceinline__ void prune_up(int p, const int32_t* __restrict__ parents,
                                         int32_t* offspring) {
  while (p >= 0) {
    const int up = parents[p];
    const int c = offspring[p];
    if (c == 1) {
      __hip_atomic_store(&offspring[p], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (atomicSub(&offspring[p], 1) != 1) {
      return;
    }
    p = up;
  }
}

// Steps 1-3 for particle i.  When the last step kept its particles (no resample), particle
// i descends from particle i and every leaf has one child.  `counts` holds each old
// particle's number of children.  Leaves are never interior nodes, so a leaf's count is
// written by its own thread only and the walks touch interior nodes only; the stores are
// agent-scope so they meet the other walks' atomics in the same place.
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_link(int N, const int32_t* __restrict__ anc, const BfScalars* sc,
                const int32_t* __restrict__ leaves, const int32_t* __restrict__ counts,
                const int32_t* __restrict__ parents, int32_t* offspring,
                int32_t* parent_of_new) {
  const int i = blockIdx.x * kTreeThreads + threadIdx.x;
  if (i >= N) return;
  const bool rs = sc->last_resampled != 0;
  const int a = rs ? anc[i] : i;
  const int c = rs ? counts[i] : 1;
  const int leaf = leaves[i];
  parent_of_new[i] = leaves[a];
  __hip_atomic_store(&offspring[leaf], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (c == 0) prune_up(parents[leaf], parents, offspring);
}

// Children counts by atomics, for ancestors without resampler counts (cnt starts at 0).
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_count_children(int N, const int32_t* __restrict__ anc, const BfScalars* sc,
                          int32_t* cnt) {
  const int i = blockIdx.x * kTreeThreads + threadIdx.x;
  if (i >= N || !sc->last_resampled) return;
  atomicAdd(&cnt[anc[i]], 1);
}

// Free slots (offspring == 0) per scan tile of `tile` slots (a multiple of kTreeThreads).
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_free_counts(int64_t M, int tile, const int32_t* __restrict__ offspring,
                       int32_t* tile_free) {
  __shared__ int red[kTreeThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * tile;
  int c = 0;
  for (int u = 0; u < tile; u += kTreeThreads) {
    const int64_t k = base + u + threadIdx.x;
    c += (k < M && offspring[k] == 0) ? 1 : 0;
  }
  // at most tile / 256 per lane: the FP32 sum is exact below 2^24 per wave
  c = (int)wave_sum_dpp((float)c);
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (lane == 0) red[w] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int k = 0; k < kTreeThreads / kWave; ++k) s += red[k];
    tile_free[blockIdx.x] = s;
  }
}

// Exclusive scan of the per-tile free counts (one block, ntile <= kTreeScanMax), the total
// in tile_pre[ntile].  Thread t owns tiles [kTreeScanPer t, kTreeScanPer (t + 1)); its
// loads are issued together.
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_tile_prefix(int ntile, const int32_t* __restrict__ tile_free, int64_t* tile_pre) {
  __shared__ int64_t wtot[kTreeThreads / kWave];
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  const int lo = t * kTreeScanPer;
  int v[kTreeScanPer];
  int64_t s = 0;
#pragma unroll
  for (int u = 0; u < kTreeScanPer; ++u) {
    const int k = min(lo + u, ntile - 1);  // in bounds; masked below
    v[u] = tile_free[k];
  }
#pragma unroll
  for (int u = 0; u < kTreeScanPer; ++u) {
    v[u] = (lo + u < ntile) ? v[u] : 0;
    s += v[u];
  }
  int64_t incl = s;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int64_t o = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += o;
  }
  if (lane == kWave - 1) wtot[w] = incl;
  __syncthreads();
  int64_t run = incl - s, total = 0;
#pragma unroll
  for (int q = 0; q < kTreeThreads / kWave; ++q) {
    if (q < w) run += wtot[q];
    total += wtot[q];
  }
#pragma unroll
  for (int u = 0; u < kTreeScanPer; ++u) {
    if (lo + u < ntile) tile_pre[lo + u] = run;
    run += v[u];
  }
  if (t == 0) tile_pre[ntile] = total;
}

// Steps 4-5: the r-th free slot k (in index order) takes new particle r, for r < N: its
// parent, its state and the particle's leaf.
template <typename T>
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_assign(int64_t M, int tile, int N, int D, const T* __restrict__ x, int Npad,
                  const int32_t* __restrict__ parent_of_new, const int64_t* __restrict__ tile_pre,
                  const int32_t* offspring, T* states, int32_t* parents, int32_t* leaves) {
  __shared__ int wsum[kTreeThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * tile;
  int64_t rank = tile_pre[blockIdx.x];
  if (rank >= N) return;  // uniform: every free slot of this tile ranks past N
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (int u = 0; u < tile && rank < N; u += kTreeThreads) {  // rank is block-uniform
    const int64_t k = base + u + threadIdx.x;
    const int f = (k < M && offspring[k] == 0) ? 1 : 0;
    const int incl = (int)wave_scan_dpp((float)f);  // at most 64: exact
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int q = 0; q < kTreeThreads / kWave; ++q) {
      if (q < w) before += wsum[q];
      total += wsum[q];
    }
    const int64_t r = rank + before + incl - f;
    if (f && r < N) {
      const int ri = (int)r;
      parents[k] = parent_of_new[ri];
      for (int d = 0; d < D; ++d) states[(size_t)d * M + k] = x[(size_t)d * Npad + ri];
      leaves[ri] = (int32_t)k;
    }
    rank += total;
    __syncthreads();
  }
}

// Path of every leaf: out[(i * (depth + 1) + t) * D + d] = x_t^{(i)}[d], t = 0 .. depth.
template <typename T>
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_paths(int N, int64_t M, int D, int depth, const T* __restrict__ states,
                 const int32_t* __restrict__ parents, const int32_t* __restrict__ leaves,
                 double* out) {
  const int i = blockIdx.x * kTreeThreads + threadIdx.x;
  if (i >= N) return;
  int node = leaves[i];
  for (int t = depth; t >= 0; --t) {
    double* o = out + ((size_t)i * (depth + 1) + t) * D;
    for (int d = 0; d < D; ++d) o[d] = (node >= 0) ? (double)states[(size_t)d * M + node] : NAN;
    node = (node >= 0) ? parents[node] : -1;
  }
}

}  // namespace ssm

using namespace ssm;

struct ssm_tree {
  ssm_bf* bf = nullptr;  // null once the filter is destroyed
  ssm_ctx* ctx = nullptr;
  int N = 0, Npad = 0, D = 0, prec = SSM_F32;
  int64_t M = 0;
  int depth = 0;
  void* states = nullptr;
  int32_t* parents = nullptr;
  int32_t* offspring = nullptr;
  int32_t* leaves = nullptr;
  int32_t* parent_of_new = nullptr;
  int32_t* cnt = nullptr;  // children per old particle (written by the resampler, or counted)
  int32_t* tile_free = nullptr;
  int64_t* tile_pre = nullptr;
  int64_t tile_cap = 0;
  // Lower bound on the free slots before the next prune: pruning only frees nodes, so
  // while it is >= N an insert needs no host read-back of the free count.
  int64_t free_lb = 0;
};

void ssm::tree_detach_filter(ssm_tree* t) { t->bf = nullptr; }

namespace {

size_t elem(const ssm_tree* t) { return t->prec == SSM_F32 ? 4 : 8; }

// Slots per scan tile: a multiple of 1024 such that at most kTreeScanMax tiles cover M.
int scan_tile(int64_t M) {
  const int64_t per = 1024 * (int64_t)kTreeScanMax;
  return (int)(1024 * ((M + per - 1) / per));
}

int64_t scan_tiles(int64_t M) {
  const int tile = scan_tile(M);
  return (M + tile - 1) / tile;
}

void tree_free(ssm_tree* t) {
  void* ps[] = {t->states, t->parents,   t->offspring, t->leaves,
                t->parent_of_new, t->cnt, t->tile_free, t->tile_pre};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

int ensure_scan(ssm_tree* t) {
  const int64_t ntile = scan_tiles(t->M);
  if (ntile <= t->tile_cap) return SSM_OK;
  if (t->tile_free) (void)hipFree(t->tile_free);
  if (t->tile_pre) (void)hipFree(t->tile_pre);
  t->tile_free = nullptr;
  t->tile_pre = nullptr;
  if (hipMalloc(&t->tile_free, sizeof(int32_t) * ntile) != hipSuccess ||
      hipMalloc(&t->tile_pre, sizeof(int64_t) * (ntile + 1)) != hipSuccess)
    return set_error(t->ctx, SSM_ERR_OOM, "path storage: scan scratch for %lld slots",
                     (long long)t->M);
  t->tile_cap = ntile;
  return SSM_OK;
}

__global__ void k_fill_i32_tree(int64_t n, int32_t* p, int32_t v) {
  const int64_t k = (int64_t)blockIdx.x * kTreeThreads + threadIdx.x;
  if (k < n) p[k] = v;
}

// Doubles the node pool (ParallelParticleTree expand!, CUDAExt.jl:213-227).
int tree_expand(ssm_tree* t) {
  ssm_ctx* ctx = t->ctx;
  hipStream_t s = ctx->stream;
  const int64_t M2 = 2 * t->M;
  if (M2 > (int64_t)INT32_MAX)
    return set_error(ctx, SSM_ERR_OOM, "path storage: more than 2^31 - 1 nodes needed");
  void* st2 = nullptr;
  int32_t *par2 = nullptr, *off2 = nullptr;
  if (hipMalloc(&st2, elem(t) * t->D * M2) != hipSuccess ||
      hipMalloc(&par2, sizeof(int32_t) * M2) != hipSuccess ||
      hipMalloc(&off2, sizeof(int32_t) * M2) != hipSuccess) {
    if (st2) (void)hipFree(st2);
    if (par2) (void)hipFree(par2);
    return set_error(ctx, SSM_ERR_OOM, "path storage: cannot grow to %lld nodes",
                     (long long)M2);
  }
  const dim3 g((unsigned)((M2 + kTreeThreads - 1) / kTreeThreads)), b(kTreeThreads);
  hipLaunchKernelGGL(k_fill_i32_tree, g, b, 0, s, M2, par2, -1);
  hipLaunchKernelGGL(k_fill_i32_tree, g, b, 0, s, M2, off2, 0);
  for (int d = 0; d < t->D; ++d)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync((char*)st2 + elem(t) * d * M2,
                                      (char*)t->states + elem(t) * d * t->M, elem(t) * t->M,
                                      hipMemcpyDeviceToDevice, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(par2, t->parents, sizeof(int32_t) * t->M,
                                    hipMemcpyDeviceToDevice, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(off2, t->offspring, sizeof(int32_t) * t->M,
                                    hipMemcpyDeviceToDevice, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  (void)hipFree(t->states);
  (void)hipFree(t->parents);
  (void)hipFree(t->offspring);
  t->states = st2;
  t->parents = par2;
  t->offspring = off2;
  t->M = M2;
  return ensure_scan(t);
}

// Free-slot counts per scan tile and their exclusive scan on the device; with `total` the
// total is read back to the host (a stream synchronisation).
int count_free(ssm_tree* t, int64_t* total) {
  ssm_ctx* ctx = t->ctx;
  hipStream_t s = ctx->stream;
  const int tile = scan_tile(t->M);
  const int64_t ntile = scan_tiles(t->M);
  hipLaunchKernelGGL(k_tree_free_counts, dim3((unsigned)ntile), dim3(kTreeThreads), 0, s, t->M,
                     tile, t->offspring, t->tile_free);
  hipLaunchKernelGGL(k_tree_tile_prefix, dim3(1), dim3(kTreeThreads), 0, s, (int)ntile,
                     t->tile_free, t->tile_pre);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  if (!total) return SSM_OK;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(total, t->tile_pre + ntile, sizeof(int64_t),
                                    hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return SSM_OK;
}

int tree_filter(ssm_tree* t, BfView* v) {
  if (!t->bf)
    return set_error(t->ctx, SSM_ERR_STATE, "path storage: its filter was destroyed");
  return bf_device_view(t->bf, v);
}

}  // namespace

extern "C" int32_t ssm_bf_tree_create(ssm_bf* bf, int64_t capacity, ssm_tree** out) {
  if (!bf || !out) return SSM_ERR_ARG;
  *out = nullptr;
  BfView v;
  int rc = bf_device_view(bf, &v);
  if (rc) return rc;
  ssm_ctx* ctx = v.ctx;
  if (capacity < v.N)
    return set_error(ctx, SSM_ERR_ARG, "path storage: capacity %lld below N = %d",
                     (long long)capacity, v.N);
  if (capacity > (int64_t)INT32_MAX)
    return set_error(ctx, SSM_ERR_ARG, "path storage: capacity above 2^31 - 1 nodes");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  ssm_tree* t = new ssm_tree();
  t->ctx = ctx;
  t->N = v.N;
  t->Npad = v.Npad;
  t->D = v.D;
  t->prec = v.prec;
  t->M = capacity;
  const size_t N = t->N;
  bool ok = hipMalloc(&t->states, elem(t) * t->D * t->M) == hipSuccess &&
            hipMalloc(&t->parents, sizeof(int32_t) * t->M) == hipSuccess &&
            hipMalloc(&t->offspring, sizeof(int32_t) * t->M) == hipSuccess &&
            hipMalloc(&t->leaves, sizeof(int32_t) * N) == hipSuccess &&
            hipMalloc(&t->parent_of_new, sizeof(int32_t) * N) == hipSuccess &&
            hipMalloc(&t->cnt, sizeof(int32_t) * N) == hipSuccess;
  if (!ok || ensure_scan(t) != SSM_OK) {
    tree_free(t);
    delete t;
    return set_error(ctx, SSM_ERR_OOM, "path storage: cannot allocate %lld nodes",
                     (long long)capacity);
  }
  rc = bf_attach_tree(bf, t, t->cnt);
  if (rc) {
    tree_free(t);
    delete t;
    return rc;
  }
  t->bf = bf;
  hipStream_t s = ctx->stream;
  const dim3 g((unsigned)((t->M + kTreeThreads - 1) / kTreeThreads)), b(kTreeThreads);
  if (t->prec == SSM_F32)
    hipLaunchKernelGGL(k_tree_roots<float>, g, b, 0, s, t->N, t->M, t->D, (const float*)v.x,
                       t->Npad, (float*)t->states, t->parents, t->offspring, t->leaves);
  else
    hipLaunchKernelGGL(k_tree_roots<double>, g, b, 0, s, t->N, t->M, t->D, (const double*)v.x,
                       t->Npad, (double*)t->states, t->parents, t->offspring, t->leaves);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  t->free_lb = t->M - t->N;
  *out = t;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_tree_insert(ssm_tree* t) {
  if (!t) return SSM_ERR_ARG;
  ssm_ctx* ctx = t->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  BfView v;
  int rc = tree_filter(t, &v);
  if (rc) return rc;
  const int N = v.N;
  hipStream_t s = ctx->stream;
  const dim3 gN((N + kTreeThreads - 1) / kTreeThreads), b(kTreeThreads);
  if (!v.counts_fresh) {
    hipLaunchKernelGGL(k_fill_i32_tree, gN, b, 0, s, (int64_t)N, t->cnt, 0);
    hipLaunchKernelGGL(k_tree_count_children, gN, b, 0, s, N, v.anc, v.sc, t->cnt);
  }
  hipLaunchKernelGGL(k_tree_link, gN, b, 0, s, N, v.anc, v.sc, t->leaves, t->cnt, t->parents,
                     t->offspring, t->parent_of_new);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  int64_t nfree = t->free_lb;
  if (nfree >= N) {
    rc = count_free(t, nullptr);
  } else {
    rc = count_free(t, &nfree);
    while (rc == SSM_OK && nfree < N) {  // ParallelParticleTree expands when short of slots
      rc = tree_expand(t);
      if (rc == SSM_OK) rc = count_free(t, &nfree);
    }
  }
  if (rc) return rc;
  const int tile = scan_tile(t->M);
  const dim3 gT((unsigned)scan_tiles(t->M));
  if (v.prec == SSM_F32)
    hipLaunchKernelGGL(k_tree_assign<float>, gT, b, 0, s, t->M, tile, N, v.D,
                       (const float*)v.x, v.Npad, t->parent_of_new, t->tile_pre, t->offspring,
                       (float*)t->states, t->parents, t->leaves);
  else
    hipLaunchKernelGGL(k_tree_assign<double>, gT, b, 0, s, t->M, tile, N, v.D,
                       (const double*)v.x, v.Npad, t->parent_of_new, t->tile_pre, t->offspring,
                       (double*)t->states, t->parents, t->leaves);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  t->free_lb = nfree - N;
  ++t->depth;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_tree_info(ssm_tree* t, int64_t* capacity, int64_t* free_slots,
                                    int32_t* depth) {
  if (!t) return SSM_ERR_ARG;
  SSM_HIP_CHECK(t->ctx, hipSetDevice(t->ctx->device));
  int64_t nfree = 0;
  const int rc = count_free(t, &nfree);
  if (rc) return rc;
  // the current leaves are childless until the next insert: they are in use, not free
  if (capacity) *capacity = t->M;
  if (free_slots) *free_slots = nfree - t->N;
  if (depth) *depth = t->depth;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_tree_ancestry(ssm_tree* t, double* paths) {
  if (!t || !paths) return SSM_ERR_ARG;
  ssm_ctx* ctx = t->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)t->N * (t->depth + 1) * t->D;
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * n) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "path storage: %zu doubles of paths", n);
  const dim3 gN((t->N + kTreeThreads - 1) / kTreeThreads), b(kTreeThreads);
  if (t->prec == SSM_F32)
    hipLaunchKernelGGL(k_tree_paths<float>, gN, b, 0, s, t->N, t->M, t->D, t->depth,
                       (const float*)t->states, t->parents, t->leaves, d);
  else
    hipLaunchKernelGGL(k_tree_paths<double>, gN, b, 0, s, t->N, t->M, t->D, t->depth,
                       (const double*)t->states, t->parents, t->leaves, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(paths, d, sizeof(double) * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess)
    return set_error(ctx, SSM_ERR_HIP, "path storage: ancestry read-back failed: %s",
                     hipGetErrorString(e));
  return SSM_OK;
}

extern "C" int32_t ssm_bf_tree_path(ssm_tree* t, int32_t i, double* path) {
  if (!t || !path) return SSM_ERR_ARG;
  ssm_ctx* ctx = t->ctx;
  if (i < 0 || i >= t->N)
    return set_error(ctx, SSM_ERR_ARG, "path storage: particle %d outside 0..%d", i, t->N - 1);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)(t->depth + 1) * t->D;
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * n) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "path storage: %zu doubles of path", n);
  // one thread walks the leaf's line upwards
  if (t->prec == SSM_F32)
    hipLaunchKernelGGL(k_tree_paths<float>, dim3(1), dim3(kTreeThreads), 0, s, 1, t->M, t->D,
                       t->depth, (const float*)t->states, t->parents, t->leaves + i, d);
  else
    hipLaunchKernelGGL(k_tree_paths<double>, dim3(1), dim3(kTreeThreads), 0, s, 1, t->M, t->D,
                       t->depth, (const double*)t->states, t->parents, t->leaves + i, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(path, d, sizeof(double) * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess)
    return set_error(ctx, SSM_ERR_HIP, "path storage: path read-back failed: %s",
                     hipGetErrorString(e));
  return SSM_OK;
}

extern "C" int32_t ssm_bf_tree_export(ssm_tree* t, int32_t* parents, int32_t* offspring,
                                      int32_t* leaves) {
  if (!t) return SSM_ERR_ARG;
  ssm_ctx* ctx = t->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (parents)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(parents, t->parents, sizeof(int32_t) * t->M,
                                      hipMemcpyDeviceToHost, s));
  if (offspring)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(offspring, t->offspring, sizeof(int32_t) * t->M,
                                      hipMemcpyDeviceToHost, s));
  if (leaves)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(leaves, t->leaves, sizeof(int32_t) * t->N,
                                      hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return SSM_OK;
}

extern "C" int32_t ssm_bf_tree_destroy(ssm_tree* t) {
  if (!t) return SSM_ERR_ARG;
  (void)hipSetDevice(t->ctx->device);
  (void)hipStreamSynchronize(t->ctx->stream);
  if (t->bf) (void)bf_attach_tree(t->bf, nullptr, nullptr);
  tree_free(t);
  delete t;
  return SSM_OK;
}
