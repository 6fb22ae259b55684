// tree.hip -- sparse path storage of a bootstrap filter's particles on the device.
//
// The device form of ParallelParticleTree (GF/ext/CUDAExt.jl:146-298), with the path
// semantics of ParticleTree (GF/src/containers.jl:420-640) and its callback
// (CUDAExt.jl:273-298): Jacob, Murray & Rubenthaler (2015), "Path storage in the particle
// filter".  Every node holds one particle state, its parent node and its live offspring
// count.  An insert of a new generation
//   1. takes each new particle's parent node as the leaf of its ancestor,
//   2. counts the children of every current leaf and stores the counts at the leaves,
//   3. prunes: a childless leaf is free, and walking up, every node whose count drops to
//      zero (one atomic decrement per child removed) is free too,
//   4. assigns the new particles to free slots in index order (the i-th new particle takes
//      the i-th free slot: an exclusive scan of the free flags), doubling the pool first
//      when fewer than N slots are free,
//   5. writes the new states, parents and leaves.
// The node pool keeps states as SoA planes states[d][M] in the filter's type.
#include <algorithm>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "internal.h"

namespace ssm {

constexpr int kTreeThreads = 256;
constexpr int kTreeTile = 1024;  // slots per block of the free-slot scan

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

// Parent node of each new particle and the children count of each current leaf.  When the
// last step kept its particles (no resample), particle i descends from particle i.
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_parents(int N, const int32_t* __restrict__ anc, const BfScalars* sc,
                   const int32_t* __restrict__ leaves, int32_t* parent_of_new, int32_t* cnt) {
  const int i = blockIdx.x * kTreeThreads + threadIdx.x;
  if (i >= N) return;
  const int a = sc->last_resampled ? anc[i] : i;
  parent_of_new[i] = leaves[a];
  atomicAdd(&cnt[a], 1);
}

__global__ void __launch_bounds__(kTreeThreads)
    k_tree_counts(int N, const int32_t* __restrict__ leaves, const int32_t* __restrict__ cnt,
                  int32_t* offspring) {
  const int j = blockIdx.x * kTreeThreads + threadIdx.x;
  if (j < N) offspring[leaves[j]] = cnt[j];
}

// Childless leaves free themselves and their now childless ancestors: the thread whose
// decrement brings a count to zero continues upwards.
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_prune(int N, const int32_t* __restrict__ leaves, const int32_t* __restrict__ cnt,
                 const int32_t* __restrict__ parents, int32_t* offspring) {
  const int j = blockIdx.x * kTreeThreads + threadIdx.x;
  if (j >= N || cnt[j] != 0) return;
  int p = parents[leaves[j]];
  while (p >= 0) {
    if (atomicSub(&offspring[p], 1) != 1) break;
    p = parents[p];
  }
}

// Free slots (offspring == 0) per block of kTreeTile slots.
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_free_counts(int64_t M, const int32_t* __restrict__ offspring, int32_t* blk_free) {
  __shared__ int red[kTreeThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kTreeTile;
  int c = 0;
  for (int u = 0; u < kTreeTile / kTreeThreads; ++u) {
    const int64_t k = base + u * kTreeThreads + threadIdx.x;
    c += (k < M && offspring[k] == 0) ? 1 : 0;
  }
  c = (int)wave_sum_dpp((float)c);  // at most 16 per lane: exact in FP32 sums up to 2^24
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (lane == 0) red[w] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int k = 0; k < kTreeThreads / kWave; ++k) s += red[k];
    blk_free[blockIdx.x] = s;
  }
}

// Exclusive scan of the per-block free counts (one block; nblk <= 256 * CH), total last.
__global__ void __launch_bounds__(256)
    k_tree_block_prefix(int nblk, const int32_t* __restrict__ blk_free, int64_t* blk_pre) {
  __shared__ int64_t part[256];
  const int t = threadIdx.x;
  const int chunk = (nblk + 255) / 256;
  const int lo = min(nblk, t * chunk), hi = min(nblk, lo + chunk);
  int64_t s = 0;
  for (int k = lo; k < hi; ++k) s += blk_free[k];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t run = 0;
    for (int k = 0; k < 256; ++k) {
      const int64_t v = part[k];
      part[k] = run;
      run += v;
    }
    blk_pre[nblk] = run;
  }
  __syncthreads();
  int64_t run = part[t];
  for (int k = lo; k < hi; ++k) {
    blk_pre[k] = run;
    run += blk_free[k];
  }
}

// The i-th free slot (in index order) goes to new particle i, for i < N.
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_assign(int64_t M, int N, const int32_t* __restrict__ offspring,
                  const int64_t* __restrict__ blk_pre, int32_t* new_leaves) {
  __shared__ int wsum[kTreeThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kTreeTile;
  int64_t rank = blk_pre[blockIdx.x];
  if (rank >= N) return;  // uniform: every free slot of this block ranks past N
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  for (int u = 0; u < kTreeTile / kTreeThreads; ++u) {
    const int64_t k = base + u * kTreeThreads + threadIdx.x;
    const int f = (k < M && offspring[k] == 0) ? 1 : 0;
    const int incl = (int)wave_scan_dpp((float)f);  // at most 64: exact
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int before = 0, total = 0;
    for (int q = 0; q < kTreeThreads / kWave; ++q) {
      if (q < w) before += wsum[q];
      total += wsum[q];
    }
    const int64_t r = rank + before + incl - f;
    if (f && r < N) new_leaves[r] = (int32_t)k;
    rank += total;
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kTreeThreads)
    k_tree_write(int N, int64_t M, int D, const T* __restrict__ x, int Npad,
                 const int32_t* __restrict__ new_leaves,
                 const int32_t* __restrict__ parent_of_new, T* states, int32_t* parents,
                 int32_t* leaves) {
  const int i = blockIdx.x * kTreeThreads + threadIdx.x;
  if (i >= N) return;
  const int node = new_leaves[i];
  parents[node] = parent_of_new[i];
  for (int d = 0; d < D; ++d) states[(size_t)d * M + node] = x[(size_t)d * Npad + i];
  leaves[i] = node;
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
  ssm_bf* bf = nullptr;
  ssm_ctx* ctx = nullptr;
  int N = 0, Npad = 0, D = 0, prec = SSM_F32;
  int64_t M = 0;
  int depth = 0;
  void* states = nullptr;
  int32_t* parents = nullptr;
  int32_t* offspring = nullptr;
  int32_t* leaves = nullptr;
  int32_t* new_leaves = nullptr;
  int32_t* parent_of_new = nullptr;
  int32_t* cnt = nullptr;
  int32_t* blk_free = nullptr;
  int64_t* blk_pre = nullptr;
  int64_t blk_cap = 0;
};

namespace {

size_t elem(const ssm_tree* t) { return t->prec == SSM_F32 ? 4 : 8; }

void tree_free(ssm_tree* t) {
  void* ps[] = {t->states, t->parents, t->offspring, t->leaves, t->new_leaves, t->parent_of_new,
                t->cnt, t->blk_free, t->blk_pre};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

int ensure_blk(ssm_tree* t) {
  const int64_t nblk = (t->M + kTreeTile - 1) / kTreeTile;
  if (nblk <= t->blk_cap) return SSM_OK;
  if (t->blk_free) (void)hipFree(t->blk_free);
  if (t->blk_pre) (void)hipFree(t->blk_pre);
  t->blk_free = nullptr;
  t->blk_pre = nullptr;
  if (hipMalloc(&t->blk_free, sizeof(int32_t) * nblk) != hipSuccess ||
      hipMalloc(&t->blk_pre, sizeof(int64_t) * (nblk + 1)) != hipSuccess)
    return set_error(t->ctx, SSM_ERR_OOM, "path storage: scan scratch for %lld slots",
                     (long long)t->M);
  t->blk_cap = nblk;
  return SSM_OK;
}

__global__ void k_fill_i32_tree(int64_t n, int32_t* p, int32_t v) {
  const int64_t k = (int64_t)blockIdx.x * kTreeThreads + threadIdx.x;
  if (k < n) p[k] = v;
}

// Doubles the node pool (ParallelParticleTree expand!, CUDAExt.jl:210-226).
int tree_expand(ssm_tree* t) {
  ssm_ctx* ctx = t->ctx;
  hipStream_t s = ctx->stream;
  const int64_t M2 = 2 * t->M;
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
  return ensure_blk(t);
}

// Free-slot counts per block and their exclusive scan; returns the total on the host.
int count_free(ssm_tree* t, int64_t* total) {
  ssm_ctx* ctx = t->ctx;
  hipStream_t s = ctx->stream;
  const int64_t nblk = (t->M + kTreeTile - 1) / kTreeTile;
  hipLaunchKernelGGL(k_tree_free_counts, dim3((unsigned)nblk), dim3(kTreeThreads), 0, s, t->M,
                     t->offspring, t->blk_free);
  hipLaunchKernelGGL(k_tree_block_prefix, dim3(1), dim3(256), 0, s, (int)nblk, t->blk_free,
                     t->blk_pre);
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(total, t->blk_pre + nblk, sizeof(int64_t),
                                    hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return SSM_OK;
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
  ssm_tree* t = new ssm_tree();
  t->ctx = ctx;
  t->N = v.N;
  t->Npad = v.Npad;
  t->D = v.D;
  t->prec = v.prec;
  const void* x = v.x;
  if (capacity > (int64_t)INT32_MAX) {
    delete t;
    return set_error(ctx, SSM_ERR_ARG, "path storage: capacity above 2^31 - 1 nodes");
  }
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  t->bf = bf;
  t->M = capacity;
  const size_t N = t->N;
  bool ok = hipMalloc(&t->states, elem(t) * t->D * t->M) == hipSuccess &&
            hipMalloc(&t->parents, sizeof(int32_t) * t->M) == hipSuccess &&
            hipMalloc(&t->offspring, sizeof(int32_t) * t->M) == hipSuccess &&
            hipMalloc(&t->leaves, sizeof(int32_t) * N) == hipSuccess &&
            hipMalloc(&t->new_leaves, sizeof(int32_t) * N) == hipSuccess &&
            hipMalloc(&t->parent_of_new, sizeof(int32_t) * N) == hipSuccess &&
            hipMalloc(&t->cnt, sizeof(int32_t) * N) == hipSuccess;
  if (!ok || ensure_blk(t) != SSM_OK) {
    tree_free(t);
    delete t;
    return set_error(ctx, SSM_ERR_OOM, "path storage: cannot allocate %lld nodes",
                     (long long)capacity);
  }
  hipStream_t s = ctx->stream;
  const dim3 g((unsigned)((t->M + kTreeThreads - 1) / kTreeThreads)), b(kTreeThreads);
  if (t->prec == SSM_F32)
    hipLaunchKernelGGL(k_tree_roots<float>, g, b, 0, s, t->N, t->M, t->D, (const float*)x,
                       t->Npad, (float*)t->states, t->parents, t->offspring, t->leaves);
  else
    hipLaunchKernelGGL(k_tree_roots<double>, g, b, 0, s, t->N, t->M, t->D, (const double*)x,
                       t->Npad, (double*)t->states, t->parents, t->offspring, t->leaves);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  *out = t;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_tree_insert(ssm_tree* t) {
  if (!t) return SSM_ERR_ARG;
  ssm_ctx* ctx = t->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  BfView v;
  int rc = bf_device_view(t->bf, &v);
  if (rc) return rc;
  const int N = v.N, Npad = v.Npad, D = v.D, prec = v.prec;
  const void* x = v.x;
  const int32_t* anc = v.anc;
  const BfScalars* sc = v.sc;
  hipStream_t s = ctx->stream;
  const dim3 gN((N + kTreeThreads - 1) / kTreeThreads), b(kTreeThreads);
  hipLaunchKernelGGL(k_fill_i32_tree, gN, b, 0, s, (int64_t)N, t->cnt, 0);
  hipLaunchKernelGGL(k_tree_parents, gN, b, 0, s, N, anc, sc, t->leaves, t->parent_of_new,
                     t->cnt);
  hipLaunchKernelGGL(k_tree_counts, gN, b, 0, s, N, t->leaves, t->cnt, t->offspring);
  hipLaunchKernelGGL(k_tree_prune, gN, b, 0, s, N, t->leaves, t->cnt, t->parents,
                     t->offspring);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  int64_t nfree = 0;
  rc = count_free(t, &nfree);
  while (rc == SSM_OK && nfree < N) {  // ParallelParticleTree expands when short of slots
    rc = tree_expand(t);
    if (rc == SSM_OK) rc = count_free(t, &nfree);
  }
  if (rc) return rc;
  const int64_t nblk = (t->M + kTreeTile - 1) / kTreeTile;
  hipLaunchKernelGGL(k_tree_assign, dim3((unsigned)nblk), dim3(kTreeThreads), 0, s, t->M, N,
                     t->offspring, t->blk_pre, t->new_leaves);
  if (prec == SSM_F32)
    hipLaunchKernelGGL(k_tree_write<float>, gN, b, 0, s, N, t->M, D, (const float*)x, Npad,
                       t->new_leaves, t->parent_of_new, (float*)t->states, t->parents,
                       t->leaves);
  else
    hipLaunchKernelGGL(k_tree_write<double>, gN, b, 0, s, N, t->M, D, (const double*)x, Npad,
                       t->new_leaves, t->parent_of_new, (double*)t->states, t->parents,
                       t->leaves);
  SSM_HIP_CHECK(ctx, hipGetLastError());
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

extern "C" int32_t ssm_bf_tree_destroy(ssm_tree* t) {
  if (!t) return SSM_ERR_ARG;
  (void)hipSetDevice(t->ctx->device);
  (void)hipStreamSynchronize(t->ctx->stream);
  tree_free(t);
  delete t;
  return SSM_OK;
}
