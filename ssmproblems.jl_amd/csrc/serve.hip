// serve.hip -- a filtering server with continuous batching on one gfx950 device.
//
// Requests are whole filter(model, algo, ys) calls (GF/src/GeneralisedFilters.jl:58-84):
// bootstrap filters (particles.jl:36-162) and Kalman filters (kalman.jl:23-79) on
// linear-Gaussian models of uneven size, particle count and horizon -- the likelihood
// evaluations an outer parameter sampler issues.  The reference has no server; this is the
// batched form of the same filters.
//
// Memory.  The server owns one device arena of the HBM budget.  A request is admitted
// when a slot is free and its device footprint (state, scratch, model, observations) can
// be carved out of the arena; it keeps that block until it finishes, so particle state
// stays device-resident for the whole run.  A request larger than the arena is refused
// at submit (SSM_ERR_ADMISSION).
//
// Bootstrap-filter lane.  All admitted filters advance together: one batched step is a
// resampler launch, a step launch and a finish launch over the tiles of every active
// filter (a tile-to-filter map routes each workgroup), so the launch count per step does
// not grow with the number of filters.  Each workgroup runs exactly the per-tile code of
// the single-filter kernels (bf_dev.h, resample_dev.h, pf_common.h), so a request's
// result does not depend on what else is in the batch.  The lane runs chunks of steps
// replayed from HIP graphs: a chunk lasts until the soonest-finishing filter is done but
// at least kMinChunk steps (a filter past its last step is skipped on the device, so it
// idles for the rest of the chunk) and at most kMaxChunk; the graph of a chunk reads the
// descriptor tables, so it is captured once per (tiles, filters, multinomial, steps) and
// replayed for every active set of that shape.  Finished filters retire and new ones are
// admitted between chunks.
//
// Kalman lane.  Waiting Kalman requests are packed into one group (ragged horizons) and
// run by the persistent size-class kernels of kf.hip on the lane's own stream, so the
// two lanes' work overlaps on the device.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "bf_dev.h"
#include "resample_dev.h"

namespace ssm {

constexpr int kServeMaxN = 1 << 18;  // one finish wave per filter (<= 256 tiles)
static_assert(kServeMaxN / kTile <= kWave * kFoldTiles, "a filter's tiles fit one fold wave");
static_assert((long long)kServeMaxN * (kMaxD + 1) < (1LL << 31),
              "the batched step's 32-bit state offsets cover every served filter");
constexpr int kMaxChunk = 64;        // batched steps between admission points
constexpr int kMinChunk = 16;        // ... at least
constexpr int kMaxGraphs = 64;       // captured chunks kept by the lane

// One active bootstrap filter as the batched kernels see it.
struct BfbFilter {
  int N, Npad, D, Dy, nb, kind, tile_base;
  int t_last;  // the filter's last step: past it every batched kernel skips the filter
  double threshold_n, logN;
  uint64_t seed;
  float* x[2];  // state planes; step t reads x[(t-1)&1] and writes x[t&1]
  float* logw;
  int32_t* anc;
  const BfModelDev<float>* model;
  const double* ys;
  BfScalars* sc;
  TileScratch ts;
  double* Wbuf;
  double* Wmid;
  double* Wroot;
};

__device__ __forceinline__ RsParams bfb_rs_params(const BfbFilter& F) {
  RsParams p;
  p.kind = F.kind;
  p.n = F.N;
  p.mode = kWeightsLogF32;
  p.w = F.logw;
  p.u = nullptr;
  p.seed = F.seed;
  p.seed_dev = &F.seed;
  p.step_host = 0;
  p.sc = F.sc;
  p.threshold_n = F.threshold_n;
  p.anc = F.anc;
  p.w_debug64 = nullptr;
  p.blk_m = F.ts.blk_m;
  p.blk_c = F.ts.blk_c;
  p.tile_sum = nullptr;
  p.prefix = F.ts.prefix;
  p.Wbuf = F.Wbuf;
  p.Wmid = F.Wmid;
  p.Wroot = F.Wroot;
  p.counts_out = nullptr;
  return p;
}

// The batched kernels run over a grid padded to a power of two of tiles (owner -1) and of
// filters (index >= *nf), so one captured chunk serves every active set of that size.
__global__ void __launch_bounds__(kTileThreads)
    k_bfb_resample(const BfbFilter* __restrict__ fs, const int32_t* __restrict__ owner) {
  const int o = owner[blockIdx.x];
  if (o < 0) return;
  const BfbFilter& F = fs[o];
  if (F.sc->step > F.t_last) return;  // finished inside this chunk
  const RsParams p = bfb_rs_params(F);
  scan_offspring_tile<kWeightsLogF32, -1, true, false>(p, blockIdx.x - F.tile_base);
}

__global__ void __launch_bounds__(kTileThreads)
    k_bfb_multinomial(const BfbFilter* __restrict__ fs, const int32_t* __restrict__ owner) {
  const int o = owner[blockIdx.x];
  if (o < 0) return;
  const BfbFilter& F = fs[o];
  if (F.kind != SSM_MULTINOMIAL || F.sc->step > F.t_last) return;
  const RsParams p = bfb_rs_params(F);
  multinomial_search_block<true>(p, blockIdx.x - F.tile_base);
}

__global__ void __launch_bounds__(kTileThreads)
    k_bfb_step(const BfbFilter* __restrict__ fs, const int32_t* __restrict__ owner) {
  const int o = owner[blockIdx.x];
  if (o < 0) return;
  const BfbFilter& F = fs[o];
  const int t = F.sc->step;
  if (t > F.t_last) return;
  StepParams<float> p;
  p.N = F.N;
  p.Npad = F.Npad;
  p.D = F.D;
  p.Dy = F.Dy;
  p.nblocks = F.nb;
  p.xin = F.x[(t - 1) & 1];
  p.xout = F.x[t & 1];
  p.logw = F.logw;
  p.anc = F.anc;
  p.model = F.model;
  p.ys = F.ys;
  p.sc = F.sc;
  p.threshold_n = F.threshold_n;
  p.logN = F.logN;
  p.seed = F.seed;
  p.seed_dev = &F.seed;
  p.noise = nullptr;
  p.partials = F.ts.partials;
  const int tile = blockIdx.x - F.tile_base;
  // the same specialisations the single-filter path launches
  if (F.D == 4 && F.Dy == 2)
    bf_step_tile<float, 4, 2, false, false, false, false, false>(p, tile);
  else if (F.D == 1 && F.Dy == 1)
    bf_step_tile<float, 1, 1, false, false, false, false, false>(p, tile);
  else
    bf_step_tile<float, 0, 0, false, false, false, false, false>(p, tile);
}

// One wave per filter: the single filter's fold in its one-wave form, which gives the same
// bits as the four-wave block for the at most 256 tiles a served filter has.
__global__ void __launch_bounds__(kWave)
    k_bfb_finish(const BfbFilter* __restrict__ fs, const int32_t* __restrict__ nf) {
  if ((int)blockIdx.x >= *nf) return;
  const BfbFilter& F = fs[blockIdx.x];
  // every thread reads the step before the block's barriers, thread 0 writes it after
  if (F.sc->step > F.t_last) return;
  pf_finish_fold4<float, kWave>(F.ts.partials, F.sc, F.nb, F.N, F.sc, F.threshold_n, F.logN,
                                F.ts, nullptr, nullptr, 1, nullptr);
}

template <int DT>
__global__ void __launch_bounds__(kTileThreads)
    k_bfb_init(int N, int Npad, int D, const BfModelDev<float>* model, float* x, float* logw,
               uint64_t seed) {
  bf_init_item<float, DT>(N, Npad, D, model, x, logw, nullptr, seed,
                          blockIdx.x * kTileThreads + threadIdx.x);
}

// First-fit allocator over the device arena (offsets, 256-byte granules), with coalescing.
class Arena {
 public:
  void reset(size_t cap) {
    cap_ = cap;
    free_.clear();
    if (cap) free_[0] = cap;
    used_ = peak_ = 0;
  }
  static size_t round(size_t b) { return (b + 255) / 256 * 256; }
  bool fits(size_t bytes) const {
    bytes = round(bytes);
    for (const auto& kv : free_)
      if (kv.second >= bytes) return true;
    return false;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = round(bytes);
    used_ -= bytes;
    auto nx = free_.lower_bound(off);
    size_t start = off, len = bytes;
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        start = pv->first;
        len += pv->second;
        free_.erase(pv);
      }
    }
    if (nx != free_.end() && off + bytes == nx->first) {
      len += nx->second;
      free_.erase(nx);
    }
    free_[start] = len;
  }
  size_t cap() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

using Clock = std::chrono::steady_clock;

struct SvReq {
  int64_t id = 0;
  int kind = SSM_REQ_BF;
  int N = 0, Npad = 0, nb = 0, T = 0, D = 0, Dy = 0, resampler = 0;
  double threshold = 1.0;
  uint64_t seed = 0;
  std::vector<double> params;  // KF: packed mu0, Sigma0, A, b, Q, H, c, R (column-major)
  std::vector<double> ys;      // T x Dy
  BfModelDev<float> model;     // BF: prepared step constants
  size_t bytes = 0, off = 0;
  int state = SSM_SERVE_QUEUED;
  int status = SSM_OK;
  double ll = 0.0;
  bool cancel = false;
  Clock::time_point t_submit, t_admit, t_done;
  // bootstrap-filter lane
  int done = 0;
  BfbFilter desc;
};

// Byte layout of one bootstrap-filter request inside its arena block.
struct BfLayout {
  size_t model, sc, x0, x1, logw, anc, rs, ys, total;
  BfLayout(int N, int Npad, int D, int Dy, int T) {
    size_t o = 0;
    auto put = [&](size_t b) {
      const size_t at = o;
      o += Arena::round(b);
      return at;
    };
    model = put(sizeof(BfModelDev<float>));
    sc = put(sizeof(BfScalars));
    x0 = put(sizeof(float) * (size_t)D * Npad);
    x1 = put(sizeof(float) * (size_t)D * Npad);
    logw = put(sizeof(float) * (size_t)Npad);
    anc = put(sizeof(int32_t) * (size_t)Npad);
    rs = put(resample_scratch_bytes(N));
    ys = put(sizeof(double) * (size_t)T * Dy);
    total = o;
  }
};

// Byte layout of one Kalman group.
struct KfLayout {
  size_t dx, dy, tf, order, poff, yoff, params, ys, ll, status, counter, total;
  KfLayout(int B, size_t nparams, size_t nys) {
    size_t o = 0;
    auto put = [&](size_t b) {
      const size_t at = o;
      o += Arena::round(b);
      return at;
    };
    dx = put(4 * (size_t)B);
    dy = put(4 * (size_t)B);
    tf = put(4 * (size_t)B);
    order = put(4 * (size_t)B);
    poff = put(8 * (size_t)B);
    yoff = put(8 * (size_t)B);
    params = put(8 * nparams);
    ys = put(8 * nys);
    ll = put(8 * (size_t)B);
    status = put(4 * (size_t)B);
    counter = put(4 * 64);
    total = o;
  }
};

inline size_t kf_param_len(int dx, int dy) {
  return 3 * (size_t)dx * dx + 2 * (size_t)dx + (size_t)dy * dx + dy + (size_t)dy * dy;
}

constexpr int kMaxKfGroup = 4096;

}  // namespace ssm

using namespace ssm;

struct ssm_server {
  ssm_ctx* user_ctx = nullptr;  // the creating context: submit errors are reported on it
  int device = 0;
  int max_active = 0;
  ssm_ctx* bf_ctx = nullptr;  // each lane has its own context (stream, error string)
  ssm_ctx* kf_ctx = nullptr;
  char* arena_base = nullptr;
  Arena arena;
  std::mutex m;
  std::condition_variable cv;       // work or memory became available
  std::condition_variable done_cv;  // a request finished
  std::unordered_map<int64_t, SvReq*> reqs;
  std::deque<int64_t> bf_queue, kf_queue;
  int64_t next_id = 1;
  bool stop = false;
  std::thread bf_thread, kf_thread;
  std::string lane_error;
  // statistics
  int64_t n_done = 0, n_failed = 0, n_cancelled = 0, bf_steps = 0, kf_groups = 0;
  int peak_active = 0;
  int64_t bf_chunks = 0, bf_captures = 0;
  double bf_capture_ms = 0.0, bf_chunk_ms = 0.0;
  // bootstrap-filter lane device tables (descriptors and tile owners) and host staging
  BfbFilter* d_desc = nullptr;
  int32_t* d_owner = nullptr;
  int32_t* d_nf = nullptr;  // active filters (the finish grid is padded beyond them)
  BfbFilter* h_desc = nullptr;
  int32_t* h_owner = nullptr;
  double* h_ll = nullptr;
  int owner_cap = 0;
  // captured chunks of batched steps, keyed by (padded tiles, padded filters, multinomial,
  // steps): the kernels read the descriptor tables, so one graph serves every active set
  // that pads to the same shape
  std::map<std::tuple<int, int, bool, int>, hipGraphExec_t> graphs;
};

namespace {

int pow2_at_least(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device side set-up of an admitted bootstrap filter (lane thread, outside the lock).
int bf_setup(ssm_server* sv, SvReq* r) {
  ssm_ctx* ctx = sv->bf_ctx;
  hipStream_t s = ctx->stream;
  char* base = sv->arena_base + r->off;
  const BfLayout L(r->N, r->Npad, r->D, r->Dy, r->T);
  BfbFilter& F = r->desc;
  F.N = r->N;
  F.Npad = r->Npad;
  F.D = r->D;
  F.Dy = r->Dy;
  F.nb = r->nb;
  F.kind = r->resampler;
  F.tile_base = 0;
  F.t_last = r->T;
  F.threshold_n = r->threshold * (double)r->N;
  F.logN = std::log((double)r->N);
  F.seed = r->seed;
  F.x[0] = (float*)(base + L.x0);
  F.x[1] = (float*)(base + L.x1);
  F.logw = (float*)(base + L.logw);
  F.anc = (int32_t*)(base + L.anc);
  F.model = (const BfModelDev<float>*)(base + L.model);
  F.ys = (const double*)(base + L.ys);
  F.sc = (BfScalars*)(base + L.sc);
  ResampleArgs a;
  resample_bind_scratch(a, base + L.rs, r->N);
  F.ts = TileScratch{a.partials, a.blk_m, a.blk_c, a.prefix, a.groups};
  F.Wbuf = a.Wbuf;
  F.Wmid = a.Wmid;
  F.Wroot = a.Wroot;
  BfScalars h;
  std::memset(&h, 0, sizeof h);
  h.s1 = (double)r->N;
  h.ess = ess_uniform(r->N);
  h.typeless = 1;
  h.step = 1;
  h.ybase = 0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(base + L.model, &r->model, sizeof r->model,
                                    hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(base + L.sc, &h, sizeof h, hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(base + L.ys, r->ys.data(), sizeof(double) * r->ys.size(),
                                    hipMemcpyHostToDevice, s));
  // the resampler scratch starts cleared, as the single-filter path's
  SSM_HIP_CHECK(ctx, hipMemsetAsync(base + L.rs, 0, resample_scratch_bytes(r->N), s));
  const dim3 g((r->Npad + kTileThreads - 1) / kTileThreads), b(kTileThreads);
  if (r->D == 4)
    hipLaunchKernelGGL(k_bfb_init<4>, g, b, 0, s, r->N, r->Npad, r->D, F.model, F.x[0], F.logw,
                       r->seed);
  else
    hipLaunchKernelGGL(k_bfb_init<0>, g, b, 0, s, r->N, r->Npad, r->D, F.model, F.x[0], F.logw,
                       r->seed);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}

void finish_req(ssm_server* sv, SvReq* r, int state, int status, double ll) {
  r->state = state;
  r->status = status;
  r->ll = ll;
  r->t_done = Clock::now();
  if (state == SSM_SERVE_DONE) ++sv->n_done;
  if (state == SSM_SERVE_FAILED) ++sv->n_failed;
  if (state == SSM_SERVE_CANCELLED) ++sv->n_cancelled;
}

void bf_lane(ssm_server* sv) {
  (void)hipSetDevice(sv->device);
  ssm_ctx* ctx = sv->bf_ctx;
  hipStream_t s = ctx->stream;
  std::vector<SvReq*> active;
  bool tables_dirty = false;
  int total_tiles = 0, tiles_pad = 0;
  bool any_multinomial = false;
  while (true) {
    std::vector<SvReq*> admitted;
    {
      std::unique_lock<std::mutex> lk(sv->m);
      auto ready = [&] {
        if (sv->stop || !active.empty()) return true;
        if (sv->bf_queue.empty()) return false;
        SvReq* h = sv->reqs[sv->bf_queue.front()];
        return h->cancel || sv->arena.fits(h->bytes);
      };
      sv->cv.wait(lk, ready);
      if (sv->stop) {
        for (SvReq* r : active) {
          sv->arena.release(r->off, r->bytes);
          finish_req(sv, r, SSM_SERVE_CANCELLED, SSM_OK, 0.0);
        }
        sv->done_cv.notify_all();
        return;
      }
      // admission: FIFO, while slots and arena memory last
      while (!sv->bf_queue.empty() &&
             (int)(active.size() + admitted.size()) < sv->max_active) {
        SvReq* r = sv->reqs[sv->bf_queue.front()];
        if (r->cancel) {
          sv->bf_queue.pop_front();
          finish_req(sv, r, SSM_SERVE_CANCELLED, SSM_OK, 0.0);
          sv->done_cv.notify_all();
          continue;
        }
        size_t off;
        if (!sv->arena.alloc(r->bytes, &off)) break;
        sv->bf_queue.pop_front();
        r->off = off;
        r->state = SSM_SERVE_RUNNING;
        r->t_admit = Clock::now();
        admitted.push_back(r);
      }
    }
    int rc = SSM_OK;
    for (SvReq* r : admitted) {
      if (rc == SSM_OK) rc = bf_setup(sv, r);
      active.push_back(r);
      tables_dirty = true;
    }
    if (rc == SSM_OK && tables_dirty && !active.empty()) {
      total_tiles = 0;
      any_multinomial = false;
      for (size_t f = 0; f < active.size(); ++f) {
        SvReq* r = active[f];
        r->desc.tile_base = total_tiles;
        sv->h_desc[f] = r->desc;
        for (int b = 0; b < r->nb; ++b) sv->h_owner[total_tiles + b] = (int32_t)f;
        total_tiles += r->nb;
        any_multinomial = any_multinomial || r->resampler == SSM_MULTINOMIAL;
      }
      tiles_pad = pow2_at_least(total_tiles);
      for (int b = total_tiles; b < tiles_pad; ++b) sv->h_owner[b] = -1;
      sv->h_owner[sv->owner_cap] = (int32_t)active.size();  // staged copy of *d_nf
      rc = hipMemcpyAsync(sv->d_desc, sv->h_desc, sizeof(BfbFilter) * active.size(),
                          hipMemcpyHostToDevice, s) == hipSuccess &&
                   hipMemcpyAsync(sv->d_owner, sv->h_owner, sizeof(int32_t) * tiles_pad,
                                  hipMemcpyHostToDevice, s) == hipSuccess &&
                   hipMemcpyAsync(sv->d_nf, sv->h_owner + sv->owner_cap, sizeof(int32_t),
                                  hipMemcpyHostToDevice, s) == hipSuccess
               ? SSM_OK
               : SSM_ERR_HIP;
      tables_dirty = false;
    }
    int k = 0;
    double chunk_capture_ms = 0.0;
    Clock::time_point chunk_start = Clock::now();
    if (rc == SSM_OK && !active.empty()) {
      // The chunk runs until the soonest-finishing filter is done, but at least kMinChunk
      // steps (filters finishing earlier skip the rest on the device) and at most
      // kMaxChunk or what the longest needs.
      int kmin = kMaxChunk, kmax = 1;
      for (SvReq* r : active) {
        kmin = std::min(kmin, r->T - r->done);
        kmax = std::max(kmax, r->T - r->done);
      }
      // chunk lengths 16, 32 or 64 (a filter past its last step idles to the chunk's end)
      k = kmin >= kMaxChunk ? kMaxChunk : kmin >= 2 * kMinChunk ? 2 * kMinChunk : kMinChunk;
      const int nf = (int)active.size();
      const int nf_pad = pow2_at_least(nf);
      const auto key = std::make_tuple(tiles_pad, nf_pad, any_multinomial, k);
      auto it = sv->graphs.find(key);
      hipGraphExec_t exec = nullptr;
      const auto tc0 = Clock::now();
      bool captured = false;
      if (it != sv->graphs.end()) {
        exec = it->second;
      } else {
        captured = true;
        if ((int)sv->graphs.size() >= kMaxGraphs) {
          for (auto& kv : sv->graphs) (void)hipGraphExecDestroy(kv.second);
          sv->graphs.clear();
        }
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        for (int step = 0; ok && step < k; ++step) {
          hipLaunchKernelGGL(k_bfb_resample, dim3(tiles_pad), dim3(kTileThreads), 0, s,
                             sv->d_desc, sv->d_owner);
          if (any_multinomial)
            hipLaunchKernelGGL(k_bfb_multinomial, dim3(tiles_pad), dim3(kTileThreads), 0, s,
                               sv->d_desc, sv->d_owner);
          hipLaunchKernelGGL(k_bfb_step, dim3(tiles_pad), dim3(kTileThreads), 0, s,
                             sv->d_desc, sv->d_owner);
          hipLaunchKernelGGL(k_bfb_finish, dim3(nf_pad), dim3(kWave), 0, s, sv->d_desc,
                             sv->d_nf);
        }
        ok = (hipStreamEndCapture(s, &g) == hipSuccess) && ok;
        ok = ok && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
        if (ok)
          sv->graphs[key] = exec;
        else
          rc = SSM_ERR_HIP;
      }
      const auto tc1 = Clock::now();
      if (rc == SSM_OK && hipGraphLaunch(exec, s) != hipSuccess) rc = SSM_ERR_HIP;
      chunk_capture_ms = captured ? ms_between(tc0, tc1) : 0.0;
      chunk_start = tc1;
      for (int f = 0; f < nf; ++f) {
        SvReq* r = active[f];
        if (r->done + k >= r->T && rc == SSM_OK)
          rc = hipMemcpyAsync(&sv->h_ll[f], &r->desc.sc->ll_total, sizeof(double),
                              hipMemcpyDeviceToHost, s) == hipSuccess
                   ? SSM_OK
                   : SSM_ERR_HIP;
      }
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        rc = SSM_ERR_HIP;
    }
    // retire finished, cancelled or (on a device error) every active filter
    const double chunk_ms = ms_between(chunk_start, Clock::now());
    std::lock_guard<std::mutex> lk(sv->m);
    sv->bf_steps += k;
    if (k > 0) {
      ++sv->bf_chunks;
      sv->bf_chunk_ms += chunk_ms;
      if (chunk_capture_ms > 0.0) {
        ++sv->bf_captures;
        sv->bf_capture_ms += chunk_capture_ms;
      }
    }
    sv->peak_active = std::max(sv->peak_active, (int)active.size());
    if (rc != SSM_OK) sv->lane_error = "bootstrap-filter lane: HIP failure";
    std::vector<SvReq*> keep;
    for (size_t f = 0; f < active.size(); ++f) {
      SvReq* r = active[f];
      r->done = std::min(r->T, r->done + k);
      if (rc != SSM_OK) {
        sv->arena.release(r->off, r->bytes);
        finish_req(sv, r, SSM_SERVE_FAILED, rc, 0.0);
      } else if (r->done >= r->T) {
        sv->arena.release(r->off, r->bytes);
        finish_req(sv, r, r->cancel ? SSM_SERVE_CANCELLED : SSM_SERVE_DONE, SSM_OK, sv->h_ll[f]);
      } else if (r->cancel) {
        sv->arena.release(r->off, r->bytes);
        finish_req(sv, r, SSM_SERVE_CANCELLED, SSM_OK, 0.0);
      } else {
        keep.push_back(r);
      }
    }
    if (keep.size() != active.size()) tables_dirty = true;
    active.swap(keep);
    sv->done_cv.notify_all();
    sv->cv.notify_all();  // memory may have been freed for the other lane
  }
}

void kf_lane(ssm_server* sv) {
  (void)hipSetDevice(sv->device);
  ssm_ctx* ctx = sv->kf_ctx;
  hipStream_t s = ctx->stream;
  while (true) {
    std::vector<SvReq*> grp;
    size_t off = 0, bytes = 0;
    size_t np = 0, ny = 0;
    {
      std::unique_lock<std::mutex> lk(sv->m);
      auto ready = [&] {
        if (sv->stop) return true;
        while (!sv->kf_queue.empty() && sv->reqs[sv->kf_queue.front()]->cancel) {
          SvReq* r = sv->reqs[sv->kf_queue.front()];
          sv->kf_queue.pop_front();
          finish_req(sv, r, SSM_SERVE_CANCELLED, SSM_OK, 0.0);
          sv->done_cv.notify_all();
        }
        if (sv->kf_queue.empty()) return false;
        const SvReq* h = sv->reqs[sv->kf_queue.front()];
        return sv->arena.fits(KfLayout(1, h->params.size(), h->ys.size()).total);
      };
      sv->cv.wait(lk, ready);
      if (sv->stop) return;
      // the longest FIFO prefix whose group fits the free arena
      int n = 0;
      size_t p_acc = 0, y_acc = 0;
      for (int64_t id : sv->kf_queue) {
        if (n >= kMaxKfGroup) break;
        const SvReq* r = sv->reqs[id];
        if (r->cancel) break;
        if (!sv->arena.fits(KfLayout(n + 1, p_acc + r->params.size(), y_acc + r->ys.size()).total))
          break;
        p_acc += r->params.size();
        y_acc += r->ys.size();
        ++n;
      }
      if (n == 0) continue;
      bytes = KfLayout(n, p_acc, y_acc).total;
      if (!sv->arena.alloc(bytes, &off)) continue;
      for (int i = 0; i < n; ++i) {
        SvReq* r = sv->reqs[sv->kf_queue.front()];
        sv->kf_queue.pop_front();
        r->state = SSM_SERVE_RUNNING;
        r->t_admit = Clock::now();
        r->off = off;
        grp.push_back(r);
      }
      np = p_acc;
      ny = y_acc;
    }
    const int B = (int)grp.size();
    const KfLayout L(B, np, ny);
    std::vector<int32_t> Dx(B), Dy(B), Tf(B), order(B);
    std::vector<int64_t> poff(B), yoff(B);
    std::vector<double> params, ys;
    params.reserve(np);
    ys.reserve(ny);
    for (int f = 0; f < B; ++f) {
      Dx[f] = grp[f]->D;
      Dy[f] = grp[f]->Dy;
      Tf[f] = grp[f]->T;
      poff[f] = (int64_t)params.size();
      yoff[f] = (int64_t)ys.size();
      params.insert(params.end(), grp[f]->params.begin(), grp[f]->params.end());
      ys.insert(ys.end(), grp[f]->ys.begin(), grp[f]->ys.end());
    }
    KfGroupDev g;
    g.B = B;
    int rc = kf_group_plan(ctx, B, Dx.data(), Dy.data(), Tf.data(), order.data(), g.count,
                           g.start);
    char* base = sv->arena_base + off;
    g.Dx = (const int32_t*)(base + L.dx);
    g.Dy = (const int32_t*)(base + L.dy);
    g.Tf = (const int32_t*)(base + L.tf);
    g.order = (const int32_t*)(base + L.order);
    g.param_off = (const int64_t*)(base + L.poff);
    g.y_off = (const int64_t*)(base + L.yoff);
    g.params = (const double*)(base + L.params);
    g.ys = (const double*)(base + L.ys);
    g.ll = (double*)(base + L.ll);
    g.status = (int32_t*)(base + L.status);
    g.counter = (uint32_t*)(base + L.counter);
    std::vector<double> ll(B);
    std::vector<int32_t> st(B);
    auto h2d = [&](const void* dst, const void* src, size_t n) {
      return hipMemcpyAsync((void*)dst, src, n, hipMemcpyHostToDevice, s) == hipSuccess;
    };
    bool ok = rc == SSM_OK && h2d(g.Dx, Dx.data(), 4 * B) && h2d(g.Dy, Dy.data(), 4 * B) &&
              h2d(g.Tf, Tf.data(), 4 * B) && h2d(g.order, order.data(), 4 * B) &&
              h2d(g.param_off, poff.data(), 8 * B) && h2d(g.y_off, yoff.data(), 8 * B) &&
              h2d(g.params, params.data(), 8 * np) && h2d(g.ys, ys.data(), 8 * ny);
    ok = ok && kf_group_launch(ctx, s, g) == SSM_OK;
    ok = ok &&
         hipMemcpyAsync(ll.data(), g.ll, 8 * B, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipMemcpyAsync(st.data(), g.status, 4 * B, hipMemcpyDeviceToHost, s) == hipSuccess;
    // wait on every path: the copies above read and write the host vectors of this scope,
    // and the arena block must be idle before it is released
    ok = (hipStreamSynchronize(s) == hipSuccess) && ok;
    std::lock_guard<std::mutex> lk(sv->m);
    sv->arena.release(off, bytes);
    ++sv->kf_groups;
    if (!ok) sv->lane_error = "Kalman lane: HIP failure";
    for (int f = 0; f < B; ++f) {
      SvReq* r = grp[f];
      if (!ok)
        finish_req(sv, r, SSM_SERVE_FAILED, SSM_ERR_HIP, 0.0);
      else if (st[f] != SSM_OK)
        finish_req(sv, r, SSM_SERVE_FAILED, SSM_ERR_NOT_PD, ll[f]);
      else
        finish_req(sv, r, r->cancel ? SSM_SERVE_CANCELLED : SSM_SERVE_DONE, SSM_OK, ll[f]);
    }
    sv->done_cv.notify_all();
    sv->cv.notify_all();
  }
}

void server_free(ssm_server* sv) {
  for (auto& kv : sv->graphs) (void)hipGraphExecDestroy(kv.second);
  sv->graphs.clear();
  if (sv->d_desc) (void)hipFree(sv->d_desc);
  if (sv->d_owner) (void)hipFree(sv->d_owner);
  if (sv->d_nf) (void)hipFree(sv->d_nf);
  if (sv->h_desc) (void)hipHostFree(sv->h_desc);
  if (sv->h_owner) (void)hipHostFree(sv->h_owner);
  if (sv->h_ll) (void)hipHostFree(sv->h_ll);
  if (sv->arena_base) (void)hipFree(sv->arena_base);
  if (sv->bf_ctx) ssm_ctx_destroy(sv->bf_ctx);
  if (sv->kf_ctx) ssm_ctx_destroy(sv->kf_ctx);
  for (auto& kv : sv->reqs) delete kv.second;
  sv->reqs.clear();
}

}  // namespace

extern "C" int32_t ssm_serve_create(ssm_ctx* ctx, uint64_t hbm_budget_bytes, int32_t max_active,
                                    ssm_server** out) {
  if (!ctx || !out) return SSM_ERR_ARG;
  *out = nullptr;
  if (hbm_budget_bytes < (1u << 20) || max_active < 1 || max_active > 4096)
    return set_error(ctx, SSM_ERR_ARG,
                     "ssm_serve_create: budget must be >= 1 MiB and 1 <= max_active <= 4096");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  ssm_server* sv = new ssm_server();
  sv->user_ctx = ctx;
  sv->device = ctx->device;
  sv->max_active = max_active;
  sv->owner_cap = pow2_at_least(max_active * (kServeMaxN / kTile));  // padded grids fit
  int rc = ssm_ctx_create(ctx->device, 0, &sv->bf_ctx);
  if (rc == SSM_OK) rc = ssm_ctx_create(ctx->device, 0, &sv->kf_ctx);
  const size_t cap = (size_t)hbm_budget_bytes / 256 * 256;
  bool ok = rc == SSM_OK && hipMalloc(&sv->arena_base, cap) == hipSuccess &&
            hipMalloc(&sv->d_desc, sizeof(BfbFilter) * max_active) == hipSuccess &&
            hipMalloc(&sv->d_owner, sizeof(int32_t) * sv->owner_cap) == hipSuccess &&
            hipMalloc(&sv->d_nf, sizeof(int32_t) * 64) == hipSuccess &&
            hipHostMalloc(&sv->h_desc, sizeof(BfbFilter) * max_active) == hipSuccess &&
            hipHostMalloc(&sv->h_owner, sizeof(int32_t) * (sv->owner_cap + 1)) == hipSuccess &&
            hipHostMalloc(&sv->h_ll, sizeof(double) * max_active) == hipSuccess;
  if (!ok) {
    server_free(sv);
    delete sv;
    return set_error(ctx, rc == SSM_OK ? SSM_ERR_OOM : rc,
                     "ssm_serve_create: cannot allocate a %llu-byte arena",
                     (unsigned long long)hbm_budget_bytes);
  }
  sv->arena.reset(cap);
  sv->bf_thread = std::thread(bf_lane, sv);
  sv->kf_thread = std::thread(kf_lane, sv);
  *out = sv;
  return SSM_OK;
}

extern "C" int32_t ssm_serve_submit(ssm_server* sv, const ssm_request* q, int64_t* id_out) {
  if (!sv || !q || !id_out) return SSM_ERR_ARG;
  ssm_ctx* ctx = sv->user_ctx;
  if (q->kind != SSM_REQ_BF && q->kind != SSM_REQ_KF)
    return set_error(ctx, SSM_ERR_ARG, "ssm_serve_submit: unknown request kind %d", q->kind);
  const ssm_lg_model* m = &q->model;
  const int dmax = q->kind == SSM_REQ_BF ? kMaxD : 64, dymax = q->kind == SSM_REQ_BF ? kMaxDy : 32;
  if (m->Dx < 1 || m->Dx > dmax || m->Dy < 1 || m->Dy > dymax)
    return set_error(ctx, SSM_ERR_ARG, "ssm_serve_submit: dims must satisfy 1<=Dx<=%d, 1<=Dy<=%d",
                     dmax, dymax);
  if (q->T < 1 || !q->ys || !m->mu0 || !m->Sigma0 || !m->A || !m->b || !m->Q || !m->H ||
      !m->c || !m->R)
    return set_error(ctx, SSM_ERR_ARG, "ssm_serve_submit: T < 1 or a null array");
  SvReq* r = new SvReq();
  r->kind = q->kind;
  r->T = q->T;
  r->D = m->Dx;
  r->Dy = m->Dy;
  r->ys.assign(q->ys, q->ys + (size_t)q->T * m->Dy);
  if (q->kind == SSM_REQ_BF) {
    if (q->N < 1 || q->N > kServeMaxN || q->resampler < SSM_SYSTEMATIC ||
        q->resampler > SSM_MULTINOMIAL || !(q->threshold >= 0.0)) {
      delete r;
      return set_error(ctx, SSM_ERR_ARG,
                       "ssm_serve_submit: need 1 <= N <= %d, a known resampler and a "
                       "threshold >= 0", kServeMaxN);
    }
    r->N = q->N;
    r->Npad = (q->N + kTile - 1) / kTile * kTile;
    r->nb = r->Npad / kTile;
    r->resampler = q->resampler;
    r->threshold = q->threshold;
    r->seed = q->seed;
    const int rc = bf_build_model<float>(ctx, m, &r->model);
    if (rc) {
      delete r;
      return rc;
    }
    r->bytes = BfLayout(r->N, r->Npad, r->D, r->Dy, r->T).total;
  } else {
    const int dx = m->Dx, dy = m->Dy;
    r->params.reserve(kf_param_len(dx, dy));
    auto put = [&](const double* p, size_t n) { r->params.insert(r->params.end(), p, p + n); };
    put(m->mu0, dx);
    put(m->Sigma0, (size_t)dx * dx);
    put(m->A, (size_t)dx * dx);
    put(m->b, dx);
    put(m->Q, (size_t)dx * dx);
    put(m->H, (size_t)dy * dx);
    put(m->c, dy);
    put(m->R, (size_t)dy * dy);
    r->bytes = KfLayout(1, r->params.size(), r->ys.size()).total;
  }
  std::lock_guard<std::mutex> lk(sv->m);
  if (r->bytes > sv->arena.cap()) {
    const size_t need = r->bytes;
    delete r;
    return set_error(ctx, SSM_ERR_ADMISSION,
                     "ssm_serve_submit: request needs %zu device bytes, the budget is %zu", need,
                     sv->arena.cap());
  }
  r->id = sv->next_id++;
  r->t_submit = Clock::now();
  sv->reqs[r->id] = r;
  (r->kind == SSM_REQ_BF ? sv->bf_queue : sv->kf_queue).push_back(r->id);
  *id_out = r->id;
  sv->cv.notify_all();
  return SSM_OK;
}

static int32_t serve_result(ssm_server* sv, int64_t id, int32_t* state, double* ll,
                            double* latency_ms, int32_t* status) {
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end()) return SSM_ERR_ARG;
  const SvReq* r = it->second;
  if (state) *state = r->state;
  if (status) *status = r->status;
  const bool fin = r->state >= SSM_SERVE_DONE;
  if (ll) *ll = fin ? r->ll : 0.0;
  if (latency_ms) *latency_ms = fin ? ms_between(r->t_submit, r->t_done) : 0.0;
  return SSM_OK;
}

extern "C" int32_t ssm_serve_poll(ssm_server* sv, int64_t id, int32_t* state, double* ll,
                                  double* latency_ms, int32_t* status) {
  if (!sv) return SSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(sv->m);
  return serve_result(sv, id, state, ll, latency_ms, status);
}

extern "C" int32_t ssm_serve_wait(ssm_server* sv, int64_t id, double timeout_ms, int32_t* state,
                                  double* ll, double* latency_ms, int32_t* status) {
  if (!sv) return SSM_ERR_ARG;
  std::unique_lock<std::mutex> lk(sv->m);
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end()) return SSM_ERR_ARG;
  const SvReq* r = it->second;
  auto fin = [&] { return r->state >= SSM_SERVE_DONE; };
  if (timeout_ms < 0)
    sv->done_cv.wait(lk, fin);
  else
    sv->done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), fin);
  return serve_result(sv, id, state, ll, latency_ms, status);
}

extern "C" int32_t ssm_serve_cancel(ssm_server* sv, int64_t id) {
  if (!sv) return SSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(sv->m);
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end()) return SSM_ERR_ARG;
  it->second->cancel = true;
  sv->cv.notify_all();
  return SSM_OK;
}

extern "C" int32_t ssm_serve_release(ssm_server* sv, int64_t id) {
  if (!sv) return SSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(sv->m);
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end() || it->second->state < SSM_SERVE_DONE) return SSM_ERR_ARG;
  delete it->second;
  sv->reqs.erase(it);
  return SSM_OK;
}

extern "C" int32_t ssm_serve_stats(ssm_server* sv, double* out, int32_t n) {
  if (!sv || !out) return SSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(sv->m);
  const double v[] = {(double)sv->n_done,       (double)sv->n_failed,
                      (double)sv->n_cancelled,  (double)sv->bf_steps,
                      (double)sv->kf_groups,    (double)sv->arena.used(),
                      (double)sv->arena.peak(), (double)sv->peak_active,
                      (double)sv->arena.cap(),  (double)(sv->bf_queue.size() + sv->kf_queue.size()),
                      (double)sv->bf_chunks,    (double)sv->bf_captures,
                      sv->bf_capture_ms,        sv->bf_chunk_ms};
  const int k = std::min<int>(n, (int)(sizeof v / sizeof v[0]));
  for (int i = 0; i < k; ++i) out[i] = v[i];
  return SSM_OK;
}

extern "C" int32_t ssm_serve_destroy(ssm_server* sv) {
  if (!sv) return SSM_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(sv->m);
    sv->stop = true;
    for (int64_t id : sv->bf_queue) finish_req(sv, sv->reqs[id], SSM_SERVE_CANCELLED, SSM_OK, 0.0);
    for (int64_t id : sv->kf_queue) finish_req(sv, sv->reqs[id], SSM_SERVE_CANCELLED, SSM_OK, 0.0);
    sv->bf_queue.clear();
    sv->kf_queue.clear();
  }
  sv->cv.notify_all();
  if (sv->bf_thread.joinable()) sv->bf_thread.join();
  if (sv->kf_thread.joinable()) sv->kf_thread.join();
  (void)hipSetDevice(sv->device);
  server_free(sv);
  delete sv;
  return SSM_OK;
}
