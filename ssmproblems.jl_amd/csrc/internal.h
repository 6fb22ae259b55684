// internal.h -- host-side internals shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/ssmhip.h"

struct ssm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t hbm_budget = 0;
  int cu_count = 0;
  std::string err;
  // Lazily grown scratch for the standalone resampler entry points.
  void* rs_scratch = nullptr;
  size_t rs_scratch_bytes = 0;
  double* rs_w = nullptr;  // FP64 staging for ssm_sample_ancestors_host
  double* rs_u = nullptr;
  int32_t* rs_a = nullptr;
  size_t rs_host_cap = 0;
  // Lazily grown device copy of the inputs of ssm_kf_loglikelihood.
  void* kfll = nullptr;
  size_t kfll_bytes = 0;
  hipEvent_t kfll_ev[2] = {nullptr, nullptr};  // around its kernel
  double kfll_ms = 0.0;
  // Lazily grown workspace of the workgroup gradient and square-root filter kernels.
  void* kfwg = nullptr;
  size_t kfwg_bytes = 0;
};

namespace ssm {

int set_error(ssm_ctx* ctx, int code, const char* fmt, ...);

#define SSM_HIP_CHECK(ctx, expr)                                                     \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return ::ssm::set_error((ctx), SSM_ERR_HIP, "%s failed: %s (%s:%d)", #expr,    \
                              hipGetErrorString(e_), __FILE__, __LINE__);            \
  } while (0)

// Grows *p to at least `bytes` of device memory (on ctx's device).
int ensure_device(ssm_ctx* ctx, void** p, size_t* cap, size_t bytes);

// ---- resampler launcher shared by the standalone entry and the particle filters ------
// Where the resampler reads its weights from.
enum WeightMode : int {
  kWeightsF64 = 0,   // explicit FP64 array
  kWeightsF32 = 1,   // explicit FP32 array
  kWeightsLogF32 = 2,  // particle-filter log-weights, FP32, normalised via BfScalars
  kWeightsLogF64 = 3,  // same, FP64
};

// FP64 scalars of a device-resident particle filter (marginalise! state and the
// resampling decision), laid out once in device memory per filter.
struct BfScalars {
  double lse;        // stored log-weights are l; normalised log-weights are l - lse
  double mxn;        // max over i of the normalised log-weights (informational)
  double s1;         // sum_i exp(l_i - max l) (informational)
  double ess;        // effective sample size of the normalised weights
  double ll_total;   // summed log-likelihood increments of the current run
  double total_w;    // resampler: sum of the weights it consumed
  int32_t typeless;  // 1 while weights are the lazy TypelessZero of initialise
  int32_t step;      // next step t (1-based) to execute
  int32_t ybase;     // observation row of step t is t - 1 - ybase
  int32_t last_resampled;
  uint32_t reserved[4];
};

struct ResampleArgs {
  int kind = SSM_SYSTEMATIC;
  int n = 0;
  int weight_mode = kWeightsF64;
  const void* w = nullptr;           // weights or log-weights
  const double* u = nullptr;         // explicit uniforms (device) or null -> Philox
  uint64_t seed = 0;
  const uint64_t* seed_dev = nullptr;  // device copy of the key (read instead of `seed`)
  uint32_t step_host = 0;            // Philox step when `sc` is null
  BfScalars* sc = nullptr;           // particle-filter scalars, or null (always resample)
  double threshold_n = 0.0;          // threshold * N (particle-filter decision)
  int32_t* anc = nullptr;            // output
  double* w_debug64 = nullptr;       // optional: weights consumed
  // tile scratch (device); particle filters pass their own, filled by their finish kernel
  double* partials = nullptr;
  double* blk_m = nullptr;
  double* blk_c = nullptr;
  double* prefix = nullptr;
  double* groups = nullptr;          // finish group records (more than 1024 tiles)
  double* Wbuf = nullptr;            // multinomial cumulative weights (tile-padded)
  double* Wmid = nullptr;            // multinomial: every 4th of them
  double* Wroot = nullptr;           // multinomial: every 64th of them
  int32_t* counts_out = nullptr;     // optional: offspring count of each particle
                                     // (systematic and stratified; ignored otherwise)
};

// Takes this device's address of the no-allocation resampler's scratch (ctx's device current).
int noalloc_setup(ssm_ctx* ctx);
// Scratch bytes the resampler needs for n weights.
size_t resample_scratch_bytes(int n);
// Carves ResampleArgs scratch pointers out of one device block of resample_scratch_bytes(n).
void resample_bind_scratch(ResampleArgs& a, void* scratch, int n);
// Enqueues the resampler kernels on `stream`.
int launch_resample(ssm_ctx* ctx, hipStream_t stream, const ResampleArgs& a);
// One categorical draw from folded log-weights (a.weight_mode a log mode, the scratch of a
// finish kernel bound in a): the first particle whose cumulative weight reaches u W_n;
// u < 0 draws u from Philox stream `philox_stream` at the step in a.sc.  Writes *out.
int launch_categorical_one(ssm_ctx* ctx, hipStream_t stream, const ResampleArgs& a, double u,
                           uint32_t philox_stream, int32_t* out);

// ---- a bootstrap filter's device buffers, for the path storage (bf.hip) -------------
struct BfView {
  ssm_ctx* ctx;
  int N, Npad, D, prec;
  const void* x;         // current state planes x[d][Npad]
  const void* logw;      // un-normalised log-weights (normalised: minus sc->lse)
  const int32_t* anc;    // ancestors of the last step (valid if sc->last_resampled)
  const BfScalars* sc;
  bool counts_fresh;     // the attached counts buffer holds the last step's offspring
                         // counts (valid if sc->last_resampled)
};
int bf_device_view(ssm_bf* bf, BfView* v);
// Attaches path storage to a filter: from then on ssm_bf_step's systematic and stratified
// resamples also write each particle's offspring count into `counts` (N entries).  One
// tree per filter; tree = null detaches.  Returns SSM_ERR_STATE when another is attached.
int bf_attach_tree(ssm_bf* bf, ssm_tree* tree, int32_t* counts);
// Called by ssm_bf_destroy: the tree loses its filter (its later calls fail cleanly).
void tree_detach_filter(ssm_tree* tree);
// Registers (on = true) or removes a dense history of the filter, so that destroying the
// filter first leaves the history's later pushes failing cleanly.
int bf_attach_history(ssm_bf* bf, ssm_history* h, bool on);
void history_detach_filter(ssm_history* h);
// A recording history takes every generation the filter's steps produce, also inside
// captured graphs.  Before steps t0 .. t0 + T - 1: room for T more generations and the
// device slot base (the push of step t writes slot t + base); *moved when the buffers were
// reallocated (graphs that hold them must go).  After the steps: the generations count.
int history_record_prepare(ssm_history* h, hipStream_t s, int t0, int T, bool* moved);
void history_record_push(ssm_history* h, hipStream_t s, const BfView& v);
void history_record_commit(ssm_history* h, int T);
// Makes h the filter's recording history (on) or stops recording into it (off).
int bf_record_history(ssm_bf* bf, ssm_history* h, bool on);
// Stored generations 0..T of a dense history (the filter's precision and layout).
struct BsHistory {
  int T;
  const void* x0;      // time-0 state planes
  const void* states;  // generation t at (t - 1) D Npad
  const double* logw;  // normalised log-weights of generation t at (t - 1) Npad
};
// Backward simulation with the filter's transition density: idx[t], t = T .. 0, drawn with
// uniform u[t]; traj (T + 1) x D.  Host outputs (either may be null).
// dyn: null (the filter's current dynamics) or T entries, dyn[t] the parameters of step
// t + 1 (only Dx, A, b and Q are read).
int bf_backward_sample(ssm_bf* bf, const BsHistory& hv, const double* u, int32_t* idx,
                       double* traj, const ssm_lg_model* dyn);

// ---- batched Kalman filters for the server (kf.hip) ---------------------------------
// Device arrays of one group of independent filters with per-filter horizons Tf; the
// parameter and observation packing is the one of ssm_kf_batch_create.
struct KfGroupDev {
  int B = 0;
  int count[3] = {0, 0, 0};  // filters per size class
  int start[3] = {0, 0, 0};  // first entry of each class in `order`
  const int32_t* Dx = nullptr;
  const int32_t* Dy = nullptr;
  const int32_t* Tf = nullptr;
  const int32_t* order = nullptr;
  const double* params = nullptr;
  const int64_t* param_off = nullptr;
  const double* ys = nullptr;
  const int64_t* y_off = nullptr;
  double* ll = nullptr;
  int32_t* status = nullptr;
  uint32_t* counter = nullptr;  // 64 words
};
// Splits filters into the size classes, costliest first (host); order has B entries.
int kf_group_plan(ssm_ctx* ctx, int B, const int32_t* Dx, const int32_t* Dy, const int32_t* Tf,
                  int32_t* order, int count[3], int start[3]);
// Enqueues the persistent class kernels of one group on stream s.
int kf_group_launch(ssm_ctx* ctx, hipStream_t s, const KfGroupDev& g);

}  // namespace ssm
