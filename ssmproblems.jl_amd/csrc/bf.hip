// bf.hip -- device-resident BootstrapFilter on a linear-Gaussian model for gfx950.
//
// Reference (paths relative to the reference checkout):
//   step       GF/src/algorithms/particles.jl:147-162  maybe_resample, then move
//   predict    particles.jl:52-74, :117-131, :228-245   x' = A x + b + chol(Q).L z
//   update     particles.jl:76-90, :133-145              logw += log N(y; H x' + c, R)
//   normalise  GF/src/containers.jl:160-173 (marginalise!), :140-141 (get_weights)
//   ESS rule   GF/src/resamplers.jl:156-183
//   filter     GF/src/GeneralisedFilters.jl:58-84
//
// Layout in HBM: state planes x[d][i] (SoA, one contiguous plane per component, padded to
// a multiple of 2048 particles), log-weights l[i], ancestors a[i] (int32), and one
// BfScalars block.  Stored log-weights are not normalised in place: the normalised value
// of marginalise! is l[i] - lse, applied when the next step reads them, so a step streams
// each array exactly once.
//
// Per step: (1) when the device-side ESS rule says so, the resampler kernel turns the
// log-weights into ancestors; (2) k_bf_step gathers x[a[i]] (or reads x[i]), propagates
// with in-register Philox noise, evaluates the observation log-density, writes x' and the
// new log-weights, and leaves one weight partial per tile of 1024 particles; (3) a one-block
// finish kernel folds the partials (FP64, fixed order) into the log-likelihood increment,
// the normaliser and the ESS that decides the next step's resampling, and prepares the
// resampler's tile prefixes (pf_common.h).  Particles are lane-striped within a tile, so
// every access of a wave is one contiguous run.  ssm_bf_run captures chunks of steps in
// HIP graphs, so T steps run without host round trips.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "bf_dev.h"

namespace ssm {

// The fused predict + update + partial-reduction kernel: one block per tile.  The pointers
// of the step's first loads lead the arguments as scalars, which the hardware preloads into
// registers (the build asks for it), so those loads need no argument load before them.
template <typename T, int DT, int DYT, bool NOISE, bool REF, bool GUIDED = false, bool APF = false>
__global__ void __launch_bounds__(kStepThreads)
    k_bf_step(const int32_t* anc, BfScalars* sc, const uint64_t* seed_dev,
              const BfModelDev<T>* model, StepParams<T> p) {
  StepParams<T> q = p;
  q.anc = anc;
  q.sc = sc;
  q.seed_dev = seed_dev;
  q.model = model;
  bf_step_tile<T, DT, DYT, NOISE, REF, GUIDED, APF>(q, blockIdx.x);
}

// k_bf_step's arguments for a parameter block
#define STEP_ARGS(p) p.anc, p.sc, p.seed_dev, p.model, p

// Conditional SMC: ancestor-sampling log-weights for the reference state x_ref[t] at the
// filter's current step t (GF/src/ancestor_sampling.jl:64-130 with the generic
// future_conditional_density): each particle's normalised log-weight (0 for the lazy
// uniform weights) plus log N(x_ref[t]; A x_i + b, Q).  One block per tile, with the
// tile's partials for their fold (thread 0 of tile 0 also hands the step to the fold's
// scalars).
template <typename T, int DT>
__global__ void __launch_bounds__(kTileThreads)
    k_as_weights(int N, int Npad, int D, const BfModelDev<T>* model, const RefModelDev<T>* rm,
                 const T* x, const T* logw, const BfScalars* sc, const T* xref, T* w_out,
                 BfScalars* as_sc, double* partials) {
  constexpr int MD = DT ? DT : kMaxD;
  if (DT) D = DT;
  const int t = sc->step;
  const bool typeless = sc->typeless;
  const T lse = (T)sc->lse;
  const int tile0 = blockIdx.x * kTile;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    as_sc->step = t;
    as_sc->typeless = 0;
  }
  // every load first (clamped in-bounds indices, no branch), then the arithmetic
  T xs[MD], xv[kTileItems][MD], lw[kTileItems];
  bool valid[kTileItems];
#pragma unroll
  for (int a = 0; a < MD; ++a) xs[a] = (a < D) ? xref[(size_t)t * D + a] : (T)0;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int i = tile0 + k * kTileThreads + threadIdx.x;
    valid[k] = i < N;
    const int ic = valid[k] ? i : N - 1;
#pragma unroll
    for (int a = 0; a < MD; ++a) xv[k][a] = (a < D) ? x[(size_t)a * Npad + ic] : (T)0;
    lw[k] = logw[ic];
  }
  T nl[kTileItems], lv[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int i = tile0 + k * kTileThreads + threadIdx.x;
    const T pre = typeless ? (T)0 : lw[k] - lse;
    const T w = pre + transition_logpdf<T, MD>(D, *model, *rm, xv[k], xs);
    if (valid[k]) w_out[i] = w;
    lv[k] = valid[k] ? w : (T)0;
    nl[k] = (T)0;
  }
  tile_partial<T>(nl, lv, valid, true, partials, gridDim.x, blockIdx.x);
}

// Auxiliary particle filter, first stage (GF/src/algorithms/particles.jl:316-337 with
// AuxiliaryResampler, GF/src/resamplers.jl:104-111): the look-ahead weights
// log w_i + log eta_i of the current particles (log w the normalised log-weights, 0 for the
// lazy uniform ones; eta the MeanPredictive score apf_log_eta at the step's observation),
// stored for the resampler, with their tile partials for the fold (thread 0 of tile 0 also
// hands the step to the fold's scalars).
template <typename T, int DT, int DYT>
__global__ void __launch_bounds__(kTileThreads)
    k_apf_weights(StepParams<T> p, T* aux_w, BfScalars* aux_sc, double* aux_partials) {
  constexpr int MD = DT ? DT : kMaxD;
  constexpr int MDY = DYT ? DYT : kMaxDy;
  const int D = DT ? DT : p.D;
  const int Dy = DYT ? DYT : p.Dy;
  const BfScalars* sc = p.sc;
  const int t = sc->step;
  const bool typeless = sc->typeless;
  const T lse = (T)sc->lse;
  const double* yrow = p.ys + (size_t)(t - 1 - sc->ybase) * Dy;
  const BfModelDev<T>* msrc = p.model;
  if constexpr (DT == 0) {
    __shared__ BfModelDev<T> lds_model;
    msrc = &model_to_lds<T>(p.model, &lds_model);
  }
  const ModelView<T, MD, MDY, (DT > 0)> mv(*msrc, D, Dy);
  T y[MDY];
#pragma unroll
  for (int q = 0; q < MDY; ++q) y[q] = (q < Dy) ? (T)yrow[q] : (T)0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    aux_sc->step = t;
    aux_sc->typeless = 0;
  }
  const int tile0 = blockIdx.x * kTile;
  T xv[kTileItems][MD], lw[kTileItems];
  bool valid[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int i = tile0 + k * kTileThreads + threadIdx.x;
    valid[k] = i < p.N;
    const int ic = valid[k] ? i : p.N - 1;
#pragma unroll
    for (int d = 0; d < MD; ++d) xv[k][d] = (d < D) ? p.xin[(size_t)d * p.Npad + ic] : (T)0;
    lw[k] = p.logw[ic];
  }
  T nl[kTileItems], lv[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int i = tile0 + k * kTileThreads + threadIdx.x;
    const T pre = typeless ? (T)0 : lw[k] - lse;
    const T w = pre + apf_log_eta<T, MD, MDY>(mv, D, Dy, xv[k], y);
    if (valid[k]) aux_w[i] = w;
    lv[k] = valid[k] ? w : (T)0;
    nl[k] = (T)0;
  }
  tile_partial<T>(nl, lv, valid, true, aux_partials, gridDim.x, blockIdx.x);
}

// CSMC-BS: backward-simulation log-weights of stored generation t (csmc.jl:418-437 with
// ancestor_weight, ancestor_sampling.jl:125-130): each particle's stored normalised
// log-weight (FP64; none for the time-0 particles, whose weights are the lazy TypelessZero)
// plus log N(x*; A x_i + b, Q), x* the state drawn at t + 1 (particle *idx_next of plane
// x_next); x_next null leaves the log-weights alone (the draw at the last time).  One block
// per tile with the tile's partials (thread 0 also marks the fold's weights explicit).
template <typename T, int DT>
__global__ void __launch_bounds__(kTileThreads)
    k_bs_weights(int N, int Npad, int D, const BfModelDev<T>* model, const RefModelDev<T>* rm,
                 const T* x, const double* logw, const T* x_next, const int32_t* idx_next,
                 T* w_out, BfScalars* as_sc, double* partials) {
  constexpr int MD = DT ? DT : kMaxD;
  if (DT) D = DT;
  const int tile0 = blockIdx.x * kTile;
  if (blockIdx.x == 0 && threadIdx.x == 0) as_sc->typeless = 0;
  // x* (the state drawn at t + 1) and every particle's operands first, then the arithmetic
  T xs[MD], xv[kTileItems][MD];
  double lw[kTileItems];
  bool valid[kTileItems];
  const int j = x_next ? *idx_next : 0;
#pragma unroll
  for (int a = 0; a < MD; ++a) xs[a] = (x_next && a < D) ? x_next[(size_t)a * Npad + j] : (T)0;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int i = tile0 + k * kTileThreads + threadIdx.x;
    valid[k] = i < N;
    const int ic = valid[k] ? i : N - 1;
#pragma unroll
    for (int a = 0; a < MD; ++a) xv[k][a] = (a < D) ? x[(size_t)a * Npad + ic] : (T)0;
    lw[k] = logw ? logw[ic] : 0.0;
  }
  T nl[kTileItems], lv[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int i = tile0 + k * kTileThreads + threadIdx.x;
    T w = (T)lw[k];
    if (x_next) w += transition_logpdf<T, MD>(D, *model, *rm, xv[k], xs);
    if (valid[k]) w_out[i] = w;
    lv[k] = valid[k] ? w : (T)0;
    nl[k] = (T)0;
  }
  tile_partial<T>(nl, lv, valid, true, partials, gridDim.x, blockIdx.x);
}

// The drawn trajectory: traj[t D + d] = generation t's state d of particle idx[t].
template <typename T>
__global__ void k_bs_gather(int Tn, int D, int Npad, const T* x0, const T* states,
                            const int32_t* idx, double* traj) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (Tn + 1) * D) return;
  const int t = e / D, d = e % D;
  const T* plane = (t == 0) ? x0 : states + (size_t)(t - 1) * D * Npad;
  traj[e] = (double)plane[(size_t)d * Npad + idx[t]];
}

// Conditional SMC: particle 0 of the initial generation is the reference's time-0 state.
template <typename T>
__global__ void k_pin_reference(int D, int Npad, T* x, const T* xref) {
  const int d = threadIdx.x;
  if (d < D) x[(size_t)d * Npad] = xref[d];
}

template <typename T, int DT>
__global__ void __launch_bounds__(kStepThreads) k_bf_init(int N, int Npad, int D,
                                                          const BfModelDev<T>* model, T* x,
                                                          T* logw, const T* noise,
                                                          uint64_t seed) {
  bf_init_item<T, DT>(N, Npad, D, model, x, logw, noise, seed,
                      blockIdx.x * kStepThreads + threadIdx.x);
}

}  // namespace ssm

using namespace ssm;

struct ssm_bf {
  ssm_ctx* ctx = nullptr;
  int N = 0, Npad = 0, D = 0, Dy = 0, prec = SSM_F32;
  uint64_t seed = 0;
  uint64_t* seed_dev = nullptr;  // device copy of the Philox key
  void* x[2] = {nullptr, nullptr};
  int cur = 0;
  void* logw = nullptr;
  int32_t* anc = nullptr;
  void* model_dev = nullptr;
  BfScalars* sc = nullptr;
  double* partials = nullptr;
  void* rs_scratch = nullptr;
  double* w_dbg = nullptr;
  double* ys = nullptr;
  size_t ys_cap = 0;
  double* ll_steps = nullptr;
  int32_t* rs_steps = nullptr;
  size_t steps_cap = 0;
  void* noise = nullptr;
  double* u = nullptr;
  bool initialised = false;
  double ess_uniform = 0.0;
  std::map<std::tuple<int, int, int, double, int>, hipGraphExec_t> graphs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int last_T = 0;
  // attached path storage: ssm_bf_step's resampler writes offspring counts for it
  ssm_tree* tree = nullptr;
  int32_t* tree_counts = nullptr;
  bool counts_fresh = false;
  std::vector<ssm_history*> histories;  // attached dense histories
  ssm_history* recording = nullptr;      // the one that takes every step's generation
  // conditional SMC: the reference trajectory and the ancestor-sampling buffers
  std::vector<double> Q_host;  // Q of the current model (column-major)
  std::vector<double> xref_host;  // the reference as given (for a model change)
  void* xref = nullptr;        // (ref_T + 1) x D, the filter's type
  size_t xref_cap = 0;         // elements
  int ref_T = -1;              // -1: no reference
  int ref_mode = SSM_CSMC_PLAIN;
  bool trans_ok = false;       // ref_model holds the current Q's constants
  int32_t* bs_idx = nullptr;   // backward simulation: drawn indices (bs_cap)
  double* bs_traj = nullptr;   // and the trajectory (bs_cap x D)
  int bs_cap = 0;
  void* ref_model = nullptr;   // RefModelDev<T>
  void* as_w = nullptr;        // Npad log-weights of the ancestor draw
  void* as_scratch = nullptr;  // their tile scratch
  BfScalars* as_sc = nullptr;  // the scalars of their fold
  int32_t* as_idx = nullptr;   // the drawn ancestor
  int32_t* draw_idx = nullptr; // ssm_bf_sample_index's result
  // guided filter: the proposal N(F x + g0 + K y, P) as given (FP64, column-major) and its
  // device constants (PropDev<T>); empty F: the bootstrap filter's latent proposal
  std::vector<double> prop_F, prop_g0, prop_K, prop_P;
  void* prop_dev = nullptr;
  // auxiliary particle filter (MeanPredictive look-ahead): its weights, tile scratch and
  // fold share the ancestor-sampling buffers (as_w, as_scratch, as_sc)
  bool apf = false;
};

namespace {

template <typename T>
int upload_model(ssm_bf* bf, const ssm_lg_model* mdl) {
  ssm_ctx* ctx = bf->ctx;
  BfModelDev<T> h;
  const int rc = bf_build_model<T>(ctx, mdl, &h);
  if (rc) return rc;
  bf->Q_host.assign(mdl->Q, mdl->Q + (size_t)mdl->Dx * mdl->Dx);
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->model_dev, &h, sizeof h, hipMemcpyHostToDevice,
                                    ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return SSM_OK;
}

int check_model(ssm_ctx* ctx, const ssm_lg_model* m) { return bf_check_model(ctx, m); }

template <typename T>
int upload_reference_t(ssm_bf* bf, int T_ref, const double* x_ref);
template <typename T>
int upload_proposal_t(ssm_bf* bf);
template <typename T>
int upload_transition_t(ssm_bf* bf);

TileScratch tile_scratch(ssm_bf* bf) {
  ResampleArgs a;
  resample_bind_scratch(a, bf->rs_scratch, bf->N);
  return TileScratch{a.partials, a.blk_m, a.blk_c, a.prefix, a.groups};
}

template <typename T>
void launch_step_t(ssm_bf* bf, hipStream_t s, double threshold_n, const T* noise) {
  StepParams<T> p;
  p.N = bf->N;
  p.Npad = bf->Npad;
  p.D = bf->D;
  p.Dy = bf->Dy;
  p.nblocks = (bf->N + kStepBlock - 1) / kStepBlock;
  p.xin = (const T*)bf->x[bf->cur];
  p.xout = (T*)bf->x[bf->cur ^ 1];
  p.logw = (T*)bf->logw;
  p.anc = bf->anc;
  p.model = (const BfModelDev<T>*)bf->model_dev;
  p.ys = bf->ys;
  p.sc = bf->sc;
  p.threshold_n = threshold_n;
  p.logN = std::log((double)bf->N);
  p.seed = bf->seed;
  p.seed_dev = bf->seed_dev;
  p.noise = noise;
  const TileScratch ts = tile_scratch(bf);
  p.partials = ts.partials;
  p.xref = (const T*)bf->xref;
  p.ref_anc = (bf->ref_mode == SSM_CSMC_ANCESTOR) ? bf->as_idx : nullptr;
  p.anc_w = bf->anc;
  p.prop = (const PropDev<T>*)bf->prop_dev;
  p.aux = bf->apf ? bf->as_sc : nullptr;
  const dim3 g(p.nblocks), b(kStepThreads);
  if (!bf->prop_F.empty()) {  // guided filter (no reference pinning)
    if (noise) {
      if (bf->D == 4 && bf->Dy == 2)
        hipLaunchKernelGGL((k_bf_step<T, 4, 2, true, false, true>), g, b, 0, s, STEP_ARGS(p));
      else if (bf->D == 1 && bf->Dy == 1)
        hipLaunchKernelGGL((k_bf_step<T, 1, 1, true, false, true>), g, b, 0, s, STEP_ARGS(p));
      else
        hipLaunchKernelGGL((k_bf_step<T, 0, 0, true, false, true>), g, b, 0, s, STEP_ARGS(p));
    } else {
      if (bf->D == 4 && bf->Dy == 2)
        hipLaunchKernelGGL((k_bf_step<T, 4, 2, false, false, true>), g, b, 0, s, STEP_ARGS(p));
      else if (bf->D == 1 && bf->Dy == 1)
        hipLaunchKernelGGL((k_bf_step<T, 1, 1, false, false, true>), g, b, 0, s, STEP_ARGS(p));
      else
        hipLaunchKernelGGL((k_bf_step<T, 0, 0, false, false, true>), g, b, 0, s, STEP_ARGS(p));
    }
    return;
  }
  if (bf->apf) {  // auxiliary particle filter (no reference, no proposal)
    if (noise) {
      if (bf->D == 4 && bf->Dy == 2)
        hipLaunchKernelGGL((k_bf_step<T, 4, 2, true, false, false, true>), g, b, 0, s, STEP_ARGS(p));
      else if (bf->D == 1 && bf->Dy == 1)
        hipLaunchKernelGGL((k_bf_step<T, 1, 1, true, false, false, true>), g, b, 0, s, STEP_ARGS(p));
      else
        hipLaunchKernelGGL((k_bf_step<T, 0, 0, true, false, false, true>), g, b, 0, s, STEP_ARGS(p));
    } else {
      if (bf->D == 4 && bf->Dy == 2)
        hipLaunchKernelGGL((k_bf_step<T, 4, 2, false, false, false, true>), g, b, 0, s, STEP_ARGS(p));
      else if (bf->D == 1 && bf->Dy == 1)
        hipLaunchKernelGGL((k_bf_step<T, 1, 1, false, false, false, true>), g, b, 0, s, STEP_ARGS(p));
      else
        hipLaunchKernelGGL((k_bf_step<T, 0, 0, false, false, false, true>), g, b, 0, s, STEP_ARGS(p));
    }
    return;
  }
  if (bf->ref_T >= 0) {  // conditional SMC: particle 0 follows the reference
    if (noise) {
      if (bf->D == 4 && bf->Dy == 2)
        hipLaunchKernelGGL((k_bf_step<T, 4, 2, true, true>), g, b, 0, s, STEP_ARGS(p));
      else
        hipLaunchKernelGGL((k_bf_step<T, 0, 0, true, true>), g, b, 0, s, STEP_ARGS(p));
    } else {
      if (bf->D == 4 && bf->Dy == 2)
        hipLaunchKernelGGL((k_bf_step<T, 4, 2, false, true>), g, b, 0, s, STEP_ARGS(p));
      else
        hipLaunchKernelGGL((k_bf_step<T, 0, 0, false, true>), g, b, 0, s, STEP_ARGS(p));
    }
    return;
  }
  if (noise) {  // parity mode: pre-drawn normals
    if (bf->D == 4 && bf->Dy == 2)
      hipLaunchKernelGGL((k_bf_step<T, 4, 2, true, false>), g, b, 0, s, STEP_ARGS(p));
    else if (bf->D == 1 && bf->Dy == 1)
      hipLaunchKernelGGL((k_bf_step<T, 1, 1, true, false>), g, b, 0, s, STEP_ARGS(p));
    else
      hipLaunchKernelGGL((k_bf_step<T, 0, 0, true, false>), g, b, 0, s, STEP_ARGS(p));
  } else {
    if (bf->D == 4 && bf->Dy == 2)
      hipLaunchKernelGGL((k_bf_step<T, 4, 2, false, false>), g, b, 0, s, STEP_ARGS(p));
    else if (bf->D == 1 && bf->Dy == 1)
      hipLaunchKernelGGL((k_bf_step<T, 1, 1, false, false>), g, b, 0, s, STEP_ARGS(p));
    else
      hipLaunchKernelGGL((k_bf_step<T, 0, 0, false, false>), g, b, 0, s, STEP_ARGS(p));
  }
}

// Enqueues the resampler kernels of one step on stream s.
int enqueue_resample(ssm_bf* bf, hipStream_t s, double threshold, int kind, const double* u_dev,
                     bool debug_weights, int32_t* counts_out = nullptr) {
  ResampleArgs a;
  a.kind = kind;
  a.n = bf->N;
  a.weight_mode = (bf->prec == SSM_F32) ? kWeightsLogF32 : kWeightsLogF64;
  // the auxiliary filter resamples on its look-ahead weights and their fold
  a.w = bf->apf ? bf->as_w : bf->logw;
  a.u = u_dev;
  a.seed = bf->seed;
  a.seed_dev = bf->seed_dev;
  a.sc = bf->apf ? bf->as_sc : bf->sc;
  a.threshold_n = threshold * (double)bf->N;
  a.anc = bf->anc;
  a.w_debug64 = debug_weights ? bf->w_dbg : nullptr;
  a.counts_out = counts_out;
  resample_bind_scratch(a, bf->apf ? bf->as_scratch : bf->rs_scratch, bf->N);
  return launch_resample(bf->ctx, s, a);
}

// Auxiliary particle filter: the look-ahead weights of the current particles and their fold
// (threshold +Inf: the fold's weights are fresh; record 0 keeps the step counter at t).
template <typename T>
void enqueue_apf_lookahead_t(ssm_bf* bf, hipStream_t s) {
  const int nb = (bf->N + kTile - 1) / kTile;
  StepParams<T> p;
  p.N = bf->N;
  p.Npad = bf->Npad;
  p.D = bf->D;
  p.Dy = bf->Dy;
  p.nblocks = nb;
  p.xin = (const T*)bf->x[bf->cur];
  p.logw = (T*)bf->logw;
  p.model = (const BfModelDev<T>*)bf->model_dev;
  p.ys = bf->ys;
  p.sc = bf->sc;
  ResampleArgs a;
  resample_bind_scratch(a, bf->as_scratch, bf->N);
  const dim3 g(nb), b(kTileThreads);
  if (bf->D == 4 && bf->Dy == 2)
    hipLaunchKernelGGL((k_apf_weights<T, 4, 2>), g, b, 0, s, p, (T*)bf->as_w, bf->as_sc,
                       a.partials);
  else if (bf->D == 1 && bf->Dy == 1)
    hipLaunchKernelGGL((k_apf_weights<T, 1, 1>), g, b, 0, s, p, (T*)bf->as_w, bf->as_sc,
                       a.partials);
  else
    hipLaunchKernelGGL((k_apf_weights<T, 0, 0>), g, b, 0, s, p, (T*)bf->as_w, bf->as_sc,
                       a.partials);
  const TileScratch ts{a.partials, a.blk_m, a.blk_c, a.prefix, a.groups};
  launch_pf_finish<T>(s, bf->N, nb, bf->as_sc, INFINITY, std::log((double)bf->N), ts, nullptr,
                      nullptr, 0);
}

// Enqueues the fused predict/update kernel of one step and flips the state buffers.
void enqueue_move(ssm_bf* bf, hipStream_t s, double threshold, const void* noise_dev) {
  const double threshold_n = threshold * (double)bf->N;
  if (bf->prec == SSM_F32)
    launch_step_t<float>(bf, s, threshold_n, (const float*)noise_dev);
  else
    launch_step_t<double>(bf, s, threshold_n, (const double*)noise_dev);
  bf->cur ^= 1;
}

// Enqueues the finish kernel (record = 1: a filter step; 0: re-derive the weight
// normalisers of imported log-weights without recording a step).
void enqueue_finish(ssm_bf* bf, hipStream_t s, double threshold, int record) {
  const int nb = (bf->N + kTile - 1) / kTile;
  const double threshold_n = threshold * (double)bf->N;
  const double logN = std::log((double)bf->N);
  const TileScratch ts = tile_scratch(bf);
  const BfScalars* aux = (bf->apf && record) ? bf->as_sc : nullptr;
  if (bf->prec == SSM_F32)
    launch_pf_finish<float>(s, bf->N, nb, bf->sc, threshold_n, logN, ts, bf->ll_steps,
                            bf->rs_steps, record, aux);
  else
    launch_pf_finish<double>(s, bf->N, nb, bf->sc, threshold_n, logN, ts, bf->ll_steps,
                             bf->rs_steps, record, aux);
}

// Conditional SMC with ancestor sampling: the reference particle's ancestor for this step,
// drawn from softmax(log w_i + log N(x_ref[t]; A x_i + b, Q)) (GF/src/algorithms/
// csmc.jl:276-292) into bf->as_idx: the weights with their tile partials, their fold and
// one categorical draw with a Philox uniform of the ancestor stream.
template <typename T>
int enqueue_ancestor_draw_t(ssm_bf* bf, hipStream_t s) {
  const int nb = (bf->N + kTile - 1) / kTile;
  const dim3 g(nb), b(kTileThreads);
  const BfModelDev<T>* md = (const BfModelDev<T>*)bf->model_dev;
  const RefModelDev<T>* rm = (const RefModelDev<T>*)bf->ref_model;
  ResampleArgs a;
  a.kind = SSM_MULTINOMIAL;
  a.n = bf->N;
  a.weight_mode = (bf->prec == SSM_F32) ? kWeightsLogF32 : kWeightsLogF64;
  a.w = bf->as_w;
  a.seed = bf->seed;
  a.seed_dev = bf->seed_dev;
  a.sc = bf->as_sc;
  resample_bind_scratch(a, bf->as_scratch, bf->N);
  if (bf->D == 4)
    hipLaunchKernelGGL((k_as_weights<T, 4>), g, b, 0, s, bf->N, bf->Npad, bf->D, md, rm,
                       (const T*)bf->x[bf->cur], (const T*)bf->logw, bf->sc,
                       (const T*)bf->xref, (T*)bf->as_w, bf->as_sc, a.partials);
  else
    hipLaunchKernelGGL((k_as_weights<T, 0>), g, b, 0, s, bf->N, bf->Npad, bf->D, md, rm,
                       (const T*)bf->x[bf->cur], (const T*)bf->logw, bf->sc,
                       (const T*)bf->xref, (T*)bf->as_w, bf->as_sc, a.partials);
  const TileScratch ts{a.partials, a.blk_m, a.blk_c, a.prefix, a.groups};
  // threshold +Inf: the fold treats the weights as fresh, as their partials were formed;
  // record 0 leaves the step counter at t (the Philox step of the draw)
  launch_pf_finish<T>(s, bf->N, nb, bf->as_sc, INFINITY, std::log((double)bf->N), ts, nullptr,
                      nullptr, 0);
  return launch_categorical_one(bf->ctx, s, a, -1.0, kStreamAncestor, bf->as_idx);
}

// Enqueues one full step (resample kernel, step kernel, finish kernel) on stream s.
int enqueue_step(ssm_bf* bf, hipStream_t s, double threshold, int kind, const double* u_dev,
                 const void* noise_dev, bool debug_weights, int32_t* counts_out = nullptr) {
  if (bf->ref_T >= 0 && bf->ref_mode == SSM_CSMC_ANCESTOR) {
    // ancestor sampling resamples at every step (csmc.jl:288)
    threshold = INFINITY;
    const int rc = (bf->prec == SSM_F32) ? enqueue_ancestor_draw_t<float>(bf, s)
                                         : enqueue_ancestor_draw_t<double>(bf, s);
    if (rc) return rc;
  }
  if (bf->apf) {  // the first stage: look-ahead weights and their fold
    if (bf->prec == SSM_F32)
      enqueue_apf_lookahead_t<float>(bf, s);
    else
      enqueue_apf_lookahead_t<double>(bf, s);
  }
  const int rc = enqueue_resample(bf, s, threshold, kind, u_dev, debug_weights, counts_out);
  if (rc) return rc;
  enqueue_move(bf, s, threshold, noise_dev);
  enqueue_finish(bf, s, threshold, 1);
  if (bf->recording) {  // DenseAncestorCallback at PostUpdate, on the device
    BfView v;
    const int rc2 = bf_device_view(bf, &v);
    if (rc2) return rc2;
    history_record_push(bf->recording, s, v);
  }
  return SSM_OK;
}

void destroy_graphs(ssm_bf* bf) {
  for (auto& kv : bf->graphs) (void)hipGraphExecDestroy(kv.second);
  bf->graphs.clear();
}

int ensure_steps(ssm_bf* bf, size_t T) {
  ssm_ctx* ctx = bf->ctx;
  if (T <= bf->steps_cap) return SSM_OK;
  destroy_graphs(bf);  // graphs hold the old buffer addresses
  (void)hipFree(bf->ys);
  (void)hipFree(bf->ll_steps);
  (void)hipFree(bf->rs_steps);
  bf->ys = nullptr;
  bf->ll_steps = nullptr;
  bf->rs_steps = nullptr;
  bf->steps_cap = 0;
  if (hipMalloc(&bf->ys, sizeof(double) * T * bf->Dy) != hipSuccess ||
      hipMalloc(&bf->ll_steps, sizeof(double) * T) != hipSuccess ||
      hipMalloc(&bf->rs_steps, sizeof(int32_t) * T) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "cannot allocate observation buffers for T=%zu", T);
  bf->steps_cap = T;
  return SSM_OK;
}

int set_step_base(ssm_bf* bf, int t, int ybase) {
  const int32_t v[2] = {t, ybase};
  SSM_HIP_CHECK(bf->ctx, hipMemcpyAsync(&bf->sc->step, v, sizeof v, hipMemcpyHostToDevice,
                                        bf->ctx->stream));
  return SSM_OK;
}

}  // namespace

extern "C" int32_t ssm_bf_create(ssm_ctx* ctx, const ssm_lg_model* model, int32_t N,
                                 int32_t precision, uint64_t seed, ssm_bf** out) {
  if (!ctx || !out) return SSM_ERR_ARG;
  *out = nullptr;
  int rc = check_model(ctx, model);
  if (rc) return rc;
  if (N < 1) return set_error(ctx, SSM_ERR_ARG, "ssm_bf_create: N must be >= 1");
  if (precision != SSM_F32 && precision != SSM_F64)
    return set_error(ctx, SSM_ERR_ARG, "ssm_bf_create: bad precision %d", precision);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  ssm_bf* bf = new ssm_bf();
  bf->ctx = ctx;
  bf->N = N;
  bf->Npad = (int)(((int64_t)N + kPadTo - 1) / kPadTo * kPadTo);
  bf->D = model->Dx;
  bf->Dy = model->Dy;
  bf->prec = precision;
  bf->seed = seed;
  const size_t es = (precision == SSM_F32) ? 4 : 8;
  bool ok = hipMalloc(&bf->x[0], es * bf->Npad * bf->D) == hipSuccess &&
            hipMalloc(&bf->x[1], es * bf->Npad * bf->D) == hipSuccess &&
            hipMalloc(&bf->logw, es * bf->Npad) == hipSuccess &&
            hipMalloc(&bf->anc, sizeof(int32_t) * bf->Npad) == hipSuccess &&
            hipMalloc(&bf->model_dev, (precision == SSM_F32) ? sizeof(BfModelDev<float>)
                                                             : sizeof(BfModelDev<double>)) ==
                hipSuccess &&
            hipMalloc(&bf->sc, sizeof(BfScalars)) == hipSuccess &&
            hipMalloc(&bf->rs_scratch, resample_scratch_bytes(N)) == hipSuccess &&
            hipMalloc(&bf->w_dbg, sizeof(double) * bf->Npad) == hipSuccess &&
            hipMalloc(&bf->seed_dev, sizeof(uint64_t)) == hipSuccess;
  if (!ok) {
    ssm_bf_destroy(bf);
    return set_error(ctx, SSM_ERR_OOM, "ssm_bf_create: allocation failed for N=%d", N);
  }
  (void)hipMemcpy(bf->seed_dev, &bf->seed, sizeof(uint64_t), hipMemcpyHostToDevice);
  (void)hipMemset(bf->x[0], 0, es * bf->Npad * bf->D);
  (void)hipMemset(bf->x[1], 0, es * bf->Npad * bf->D);
  (void)hipMemset(bf->logw, 0, es * bf->Npad);
  (void)hipMemset(bf->anc, 0, sizeof(int32_t) * bf->Npad);
  (void)hipMemset(bf->sc, 0, sizeof(BfScalars));
  (void)hipMemset(bf->rs_scratch, 0, resample_scratch_bytes(N));
  if (hipEventCreate(&bf->ev0) != hipSuccess || hipEventCreate(&bf->ev1) != hipSuccess) {
    ssm_bf_destroy(bf);
    return set_error(ctx, SSM_ERR_HIP, "ssm_bf_create: event creation failed");
  }
  rc = (precision == SSM_F32) ? upload_model<float>(bf, model) : upload_model<double>(bf, model);
  if (rc) {
    ssm_bf_destroy(bf);
    return rc;
  }
  rc = ensure_steps(bf, 1);
  if (rc) {
    ssm_bf_destroy(bf);
    return rc;
  }
  bf->ess_uniform = ess_uniform(N);
  *out = bf;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_destroy(ssm_bf* bf) {
  if (!bf) return SSM_ERR_ARG;
  (void)hipSetDevice(bf->ctx->device);
  (void)hipStreamSynchronize(bf->ctx->stream);
  destroy_graphs(bf);
  if (bf->tree) tree_detach_filter(bf->tree);
  for (ssm_history* h : bf->histories) history_detach_filter(h);
  void* ptrs[] = {bf->x[0],      bf->x[1],     bf->logw,     bf->anc,   bf->model_dev,
                  bf->sc,        bf->partials, bf->rs_scratch, bf->w_dbg, bf->ys,
                  bf->ll_steps,  bf->rs_steps, bf->noise,    bf->u,     bf->xref,
                  bf->ref_model, bf->as_w,     bf->as_scratch, bf->as_sc, bf->as_idx,
                  bf->draw_idx,  bf->seed_dev, bf->bs_idx,   bf->bs_traj, bf->prop_dev};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (bf->ev0) (void)hipEventDestroy(bf->ev0);
  if (bf->ev1) (void)hipEventDestroy(bf->ev1);
  delete bf;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_set_model(ssm_bf* bf, const ssm_lg_model* model) {
  if (!bf) return SSM_ERR_ARG;
  int rc = check_model(bf->ctx, model);
  if (rc) return rc;
  if (model->Dx != bf->D || model->Dy != bf->Dy)
    return set_error(bf->ctx, SSM_ERR_ARG, "ssm_bf_set_model: dimensions changed");
  rc = (bf->prec == SSM_F32) ? upload_model<float>(bf, model) : upload_model<double>(bf, model);
  bf->trans_ok = false;
  if (!rc && !bf->prop_F.empty())  // the proposal's weight constants follow Q
    rc = (bf->prec == SSM_F32) ? upload_proposal_t<float>(bf) : upload_proposal_t<double>(bf);
  if (rc || bf->ref_T < 0) return rc;
  // the reference's transition constants follow Q
  return (bf->prec == SSM_F32) ? upload_reference_t<float>(bf, bf->ref_T, bf->xref_host.data())
                               : upload_reference_t<double>(bf, bf->ref_T, bf->xref_host.data());
}

template <typename T>
static int upload_noise(ssm_bf* bf, const double* z) {
  ssm_ctx* ctx = bf->ctx;
  if (!bf->noise) {
    if (hipMalloc(&bf->noise, sizeof(T) * bf->Npad * bf->D) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "noise buffer");
  }
  std::vector<T> h((size_t)bf->Npad * bf->D, (T)0);
  for (int i = 0; i < bf->N; ++i)
    for (int d = 0; d < bf->D; ++d) h[(size_t)d * bf->Npad + i] = (T)z[(size_t)i * bf->D + d];
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->noise, h.data(), sizeof(T) * h.size(),
                                    hipMemcpyHostToDevice, ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return SSM_OK;
}

extern "C" int32_t ssm_bf_initialise(ssm_bf* bf, const double* z) {
  if (!bf) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  bf->counts_fresh = false;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (z) {
    int rc = (bf->prec == SSM_F32) ? upload_noise<float>(bf, z) : upload_noise<double>(bf, z);
    if (rc) return rc;
  }
  BfScalars h;
  std::memset(&h, 0, sizeof h);
  h.lse = 0.0;
  h.mxn = 0.0;
  h.s1 = (double)bf->N;
  h.ess = bf->ess_uniform;
  h.typeless = 1;
  h.step = 1;
  h.ybase = 0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->sc, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
  const dim3 g((bf->Npad + kStepThreads - 1) / kStepThreads), b(kStepThreads);
  bf->cur = 0;
  if (bf->prec == SSM_F32) {
    const float* nz = z ? (const float*)bf->noise : nullptr;
    if (bf->D == 4)
      hipLaunchKernelGGL((k_bf_init<float, 4>), g, b, 0, ctx->stream, bf->N, bf->Npad, bf->D,
                         (const BfModelDev<float>*)bf->model_dev, (float*)bf->x[0],
                         (float*)bf->logw, nz, bf->seed);
    else
      hipLaunchKernelGGL((k_bf_init<float, 0>), g, b, 0, ctx->stream, bf->N, bf->Npad, bf->D,
                         (const BfModelDev<float>*)bf->model_dev, (float*)bf->x[0],
                         (float*)bf->logw, nz, bf->seed);
  } else {
    const double* nz = z ? (const double*)bf->noise : nullptr;
    hipLaunchKernelGGL((k_bf_init<double, 0>), g, b, 0, ctx->stream, bf->N, bf->Npad, bf->D,
                       (const BfModelDev<double>*)bf->model_dev, (double*)bf->x[0],
                       (double*)bf->logw, nz, bf->seed);
  }
  if (bf->ref_T >= 0) {  // conditional SMC: particle 0 starts at the reference (particles.jl:40-50)
    if (bf->prec == SSM_F32)
      hipLaunchKernelGGL(k_pin_reference<float>, dim3(1), dim3(64), 0, ctx->stream, bf->D,
                         bf->Npad, (float*)bf->x[0], (const float*)bf->xref);
    else
      hipLaunchKernelGGL(k_pin_reference<double>, dim3(1), dim3(64), 0, ctx->stream, bf->D,
                         bf->Npad, (double*)bf->x[0], (const double*)bf->xref);
  }
  SSM_HIP_CHECK(ctx, hipGetLastError());
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  bf->initialised = true;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_step(ssm_bf* bf, int32_t t, const double* y, double threshold,
                               int32_t resampler, const double* u, const double* z,
                               double* ll_inc, int32_t* did_resample) {
  if (!bf) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  if (!bf->initialised) return set_error(ctx, SSM_ERR_STATE, "ssm_bf_step before initialise");
  if (!y || t < 1) return set_error(ctx, SSM_ERR_ARG, "ssm_bf_step: bad t or null y");
  if (bf->ref_T >= 0 && t > bf->ref_T)
    return set_error(ctx, SSM_ERR_ARG, "ssm_bf_step: t = %d past the reference's %d steps", t,
                     bf->ref_T);
  if (resampler < SSM_SYSTEMATIC || resampler > SSM_MULTINOMIAL)
    return set_error(ctx, SSM_ERR_ARG, "ssm_bf_step: unknown resampler %d", resampler);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int rc;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->ys, y, sizeof(double) * bf->Dy, hipMemcpyHostToDevice,
                                    ctx->stream));
  rc = set_step_base(bf, t, t - 1);
  if (rc) return rc;
  const double* u_dev = nullptr;
  if (u) {
    if (!bf->u && hipMalloc(&bf->u, sizeof(double) * bf->Npad) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "uniform buffer");
    const size_t nu =
        (resampler == SSM_SYSTEMATIC || resampler == SSM_SYSTEMATIC_MURRAY) ? 1 : bf->N;
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->u, u, sizeof(double) * nu, hipMemcpyHostToDevice,
                                      ctx->stream));
    u_dev = bf->u;
  }
  const void* noise_dev = nullptr;
  if (z) {
    rc = (bf->prec == SSM_F32) ? upload_noise<float>(bf, z) : upload_noise<double>(bf, z);
    if (rc) return rc;
    noise_dev = bf->noise;
  }
  if (bf->recording) {
    bool moved = false;
    rc = history_record_prepare(bf->recording, ctx->stream, t, 1, &moved);
    if (rc) return rc;
    if (moved) destroy_graphs(bf);
  }
  bf->counts_fresh = false;
  // with a reference, particle 0's ancestor is replaced after the resampler counted
  // offspring, so the tree recounts from the ancestors instead
  const bool pass_counts = bf->tree_counts != nullptr && bf->ref_T < 0;
  rc = enqueue_step(bf, ctx->stream, threshold, resampler, u_dev, noise_dev, true,
                    pass_counts ? bf->tree_counts : nullptr);
  if (rc) return rc;
  SSM_HIP_CHECK(ctx, hipGetLastError());
  double ll = 0.0;
  int32_t did = 0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&ll, bf->ll_steps, sizeof ll, hipMemcpyDeviceToHost,
                                    ctx->stream));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&did, bf->rs_steps, sizeof did, hipMemcpyDeviceToHost,
                                    ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  bf->counts_fresh = pass_counts && resampler != SSM_MULTINOMIAL;
  if (bf->recording) history_record_commit(bf->recording, 1);
  if (ll_inc) *ll_inc = ll;
  if (did_resample) *did_resample = did;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_run(ssm_bf* bf, int32_t t0, int32_t T, const double* ys,
                              double threshold, int32_t resampler, double* ll_total,
                              double* ll_steps) {
  if (!bf) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  if (!bf->initialised) return set_error(ctx, SSM_ERR_STATE, "ssm_bf_run before initialise");
  if (T < 1 || t0 < 1 || !ys) return set_error(ctx, SSM_ERR_ARG, "ssm_bf_run: bad arguments");
  if (bf->ref_T >= 0 && t0 + T - 1 > bf->ref_T)
    return set_error(ctx, SSM_ERR_ARG, "ssm_bf_run: steps past the reference's %d", bf->ref_T);
  bf->counts_fresh = false;
  if (resampler < SSM_SYSTEMATIC || resampler > SSM_MULTINOMIAL)
    return set_error(ctx, SSM_ERR_ARG, "ssm_bf_run: unknown resampler %d", resampler);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int rc = ensure_steps(bf, (size_t)T);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  if (bf->recording) {
    bool moved = false;
    rc = history_record_prepare(bf->recording, s, t0, T, &moved);
    if (rc) return rc;
    if (moved) destroy_graphs(bf);
  }
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->ys, ys, sizeof(double) * T * bf->Dy,
                                    hipMemcpyHostToDevice, s));
  const double zero = 0.0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&bf->sc->ll_total, &zero, sizeof zero,
                                    hipMemcpyHostToDevice, s));
  rc = set_step_base(bf, t0, t0 - 1);
  if (rc) return rc;
  constexpr int kChunk = 16;  // even: the x ping-pong returns to its start
  const int nfull = T / kChunk, rem = T % kChunk;
  auto one_step = [&]() {
    return enqueue_step(bf, s, threshold, resampler, nullptr, nullptr, false);
  };
  hipGraphExec_t exec = nullptr;
  if (nfull > 0) {
    const auto key = std::make_tuple(kChunk, bf->cur, resampler, threshold,
                                     bf->ref_T >= 0 ? 1 + bf->ref_mode : 0);
    auto it = bf->graphs.find(key);
    if (it == bf->graphs.end()) {
      hipGraph_t g;
      SSM_HIP_CHECK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int cur0 = bf->cur;
      for (int k = 0; k < kChunk && rc == SSM_OK; ++k) rc = one_step();
      bf->cur = cur0;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (rc) return rc;
      if (e != hipSuccess)
        return set_error(ctx, SSM_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess)
        return set_error(ctx, SSM_ERR_HIP, "graph instantiate failed: %s",
                         hipGetErrorString(e));
      bf->graphs[key] = exec;
    } else {
      exec = it->second;
    }
  }
  SSM_HIP_CHECK(ctx, hipEventRecord(bf->ev0, s));
  for (int k = 0; k < nfull; ++k) SSM_HIP_CHECK(ctx, hipGraphLaunch(exec, s));
  for (int k = 0; k < rem && rc == SSM_OK; ++k) rc = one_step();
  if (rc) return rc;
  SSM_HIP_CHECK(ctx, hipEventRecord(bf->ev1, s));
  SSM_HIP_CHECK(ctx, hipGetLastError());
  double tot = 0.0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&tot, &bf->sc->ll_total, sizeof tot,
                                    hipMemcpyDeviceToHost, s));
  if (ll_steps)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(ll_steps, bf->ll_steps, sizeof(double) * T,
                                      hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  float ms = 0.f;
  SSM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, bf->ev0, bf->ev1));
  bf->last_ms = ms;
  bf->last_T = T;
  if (bf->recording) history_record_commit(bf->recording, T);
  if (ll_total) *ll_total = tot;
  return SSM_OK;
}

template <typename T>
static int export_t(ssm_bf* bf, double* x, double* logw, int32_t* anc, double* w) {
  ssm_ctx* ctx = bf->ctx;
  hipStream_t s = ctx->stream;
  BfScalars h;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&h, bf->sc, sizeof h, hipMemcpyDeviceToHost, s));
  std::vector<T> xb, lb;
  if (x) {
    xb.resize((size_t)bf->Npad * bf->D);
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(xb.data(), bf->x[bf->cur], sizeof(T) * xb.size(),
                                      hipMemcpyDeviceToHost, s));
  }
  if (logw) {
    lb.resize(bf->N);
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(lb.data(), bf->logw, sizeof(T) * bf->N,
                                      hipMemcpyDeviceToHost, s));
  }
  if (anc)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(anc, bf->anc, sizeof(int32_t) * bf->N,
                                      hipMemcpyDeviceToHost, s));
  if (w)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(w, bf->w_dbg, sizeof(double) * bf->N,
                                      hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (x)
    for (int i = 0; i < bf->N; ++i)
      for (int d = 0; d < bf->D; ++d)
        x[(size_t)i * bf->D + d] = (double)xb[(size_t)d * bf->Npad + i];
  if (logw) {
    const T lse = (T)h.lse;
    for (int i = 0; i < bf->N; ++i) logw[i] = h.typeless ? 0.0 : (double)(T)(lb[i] - lse);
  }
  if (anc && !h.last_resampled)
    for (int i = 0; i < bf->N; ++i) anc[i] = i;  // preserve_sample: a_i = i
  return SSM_OK;
}

extern "C" int32_t ssm_bf_export(ssm_bf* bf, double* x, double* logw, int32_t* ancestors,
                                 double* w) {
  if (!bf) return SSM_ERR_ARG;
  SSM_HIP_CHECK(bf->ctx, hipSetDevice(bf->ctx->device));
  return (bf->prec == SSM_F32) ? export_t<float>(bf, x, logw, ancestors, w)
                               : export_t<double>(bf, x, logw, ancestors, w);
}

template <typename T>
static int import_t(ssm_bf* bf, const double* x, const double* logw) {
  ssm_ctx* ctx = bf->ctx;
  hipStream_t s = ctx->stream;
  std::vector<T> xb((size_t)bf->Npad * bf->D, (T)0), lb(bf->Npad, (T)0);
  for (int i = 0; i < bf->N; ++i)
    for (int d = 0; d < bf->D; ++d) xb[(size_t)d * bf->Npad + i] = (T)x[(size_t)i * bf->D + d];
  double mx = -INFINITY;
  for (int i = 0; i < bf->N; ++i) {
    lb[i] = (T)logw[i];
    mx = std::max(mx, (double)lb[i]);
  }
  double s1 = 0.0, s2 = 0.0;
  for (int i = 0; i < bf->N; ++i) {
    const double e = std::exp((double)lb[i] - mx);
    s1 += e;
    s2 += e * e;
  }
  BfScalars h;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&h, bf->sc, sizeof h, hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  h.lse = 0.0;
  h.mxn = mx;
  h.s1 = s1;
  h.ess = std::min((s1 * s1) / s2, (double)bf->N);
  h.typeless = 0;
  h.last_resampled = 0;
  bf->cur = 0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->x[0], xb.data(), sizeof(T) * xb.size(),
                                    hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->logw, lb.data(), sizeof(T) * lb.size(),
                                    hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->sc, &h, sizeof h, hipMemcpyHostToDevice, s));
  // tile partials of the imported log-weights, then the normalisers the resampler uses
  const int nb = (bf->N + kTile - 1) / kTile;
  const TileScratch ts = tile_scratch(bf);
  hipLaunchKernelGGL(k_tile_partials<T>, dim3(nb), dim3(kTileThreads), 0, s, bf->N,
                     (const T*)bf->logw, (const T*)nullptr, bf->sc, -1.0, ts.partials);
  enqueue_finish(bf, s, 0.0, 0);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  bf->initialised = true;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_import(ssm_bf* bf, const double* x, const double* logw) {
  if (!bf || !x || !logw) return bf ? set_error(bf->ctx, SSM_ERR_ARG, "null input") : SSM_ERR_ARG;
  SSM_HIP_CHECK(bf->ctx, hipSetDevice(bf->ctx->device));
  bf->counts_fresh = false;
  return (bf->prec == SSM_F32) ? import_t<float>(bf, x, logw) : import_t<double>(bf, x, logw);
}

extern "C" int32_t ssm_bf_profile(ssm_bf* bf, int32_t t0, int32_t T, const double* ys,
                                  double threshold, int32_t resampler, int32_t reps,
                                  double* ms_kernels, double* ll_total) {
  if (bf && bf->apf)
    return set_error(bf->ctx, SSM_ERR_ARG, "ssm_bf_profile: times the bootstrap filter only");
  if (!bf) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  if (!bf->initialised) return set_error(ctx, SSM_ERR_STATE, "ssm_bf_profile before initialise");
  if (T < 1 || t0 < 1 || !ys || reps < 0)
    return set_error(ctx, SSM_ERR_ARG, "ssm_bf_profile: bad arguments");
  if (resampler < SSM_SYSTEMATIC || resampler > SSM_MULTINOMIAL)
    return set_error(ctx, SSM_ERR_ARG, "ssm_bf_profile: unknown resampler %d", resampler);
  if (bf->ref_T >= 0)
    return set_error(ctx, SSM_ERR_STATE, "ssm_bf_profile: times the unconditional filter only");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  bf->counts_fresh = false;
  int rc = ensure_steps(bf, (size_t)T);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->ys, ys, sizeof(double) * T * bf->Dy,
                                    hipMemcpyHostToDevice, s));
  const double zero = 0.0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&bf->sc->ll_total, &zero, sizeof zero,
                                    hipMemcpyHostToDevice, s));
  rc = set_step_base(bf, t0, t0 - 1);
  if (rc) return rc;
  std::vector<hipEvent_t> ev(4 * (size_t)T, nullptr);
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) return set_error(ctx, SSM_ERR_HIP, "event create");
  for (int k = 0; k < T && rc == SSM_OK; ++k) {
    (void)hipEventRecord(ev[4 * k], s);
    rc = enqueue_resample(bf, s, threshold, resampler, nullptr, false);
    (void)hipEventRecord(ev[4 * k + 1], s);
    enqueue_move(bf, s, threshold, nullptr);
    (void)hipEventRecord(ev[4 * k + 2], s);
    enqueue_finish(bf, s, threshold, 1);
    (void)hipEventRecord(ev[4 * k + 3], s);
  }
  double tot = 0.0;
  if (rc == SSM_OK) {
    (void)hipMemcpyAsync(&tot, &bf->sc->ll_total, sizeof tot, hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
  }
  // back-to-back launches on the final state; the finish kernel runs with record = 0 so
  // the step counter stays, and the step kernel is launched with the same inputs each time
  double b2b[3] = {0.0, 0.0, 0.0};
  // step counter back to the last step, so the step kernel reads observation row T - 1
  if (rc == SSM_OK && reps > 0) rc = set_step_base(bf, t0 + T - 1, t0 - 1);
  if (rc == SSM_OK && reps > 0) {
    hipEvent_t e[5];
    for (auto& x : e)
      if (hipEventCreate(&x) != hipSuccess) return set_error(ctx, SSM_ERR_HIP, "event create");
    (void)hipEventRecord(e[0], s);
    for (int k = 0; k < reps; ++k) enqueue_move(bf, s, threshold, nullptr);
    (void)hipEventRecord(e[1], s);
    for (int k = 0; k < reps && rc == SSM_OK; ++k)
      rc = enqueue_resample(bf, s, threshold, resampler, nullptr, false);
    (void)hipEventRecord(e[2], s);
    for (int k = 0; k < reps; ++k) enqueue_finish(bf, s, threshold, 0);
    (void)hipEventRecord(e[3], s);
    (void)hipStreamSynchronize(s);
    float m01 = 0.f, m12 = 0.f, m23 = 0.f;
    (void)hipEventElapsedTime(&m01, e[0], e[1]);
    (void)hipEventElapsedTime(&m12, e[1], e[2]);
    (void)hipEventElapsedTime(&m23, e[2], e[3]);
    b2b[1] = m01 / reps;
    b2b[0] = m12 / reps;
    b2b[2] = m23 / reps;
    for (auto& x : e) (void)hipEventDestroy(x);
  }
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < T && rc == SSM_OK; ++k)
    for (int j = 0; j < 3; ++j) {
      float m = 0.f;
      (void)hipEventElapsedTime(&m, ev[4 * k + j], ev[4 * k + j + 1]);
      acc[j] += m;
    }
  if (rc) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
  }
  SSM_HIP_CHECK(ctx, hipGetLastError());
  if (ms_kernels)
    for (int j = 0; j < 3; ++j) {
      ms_kernels[j] = acc[j] / T;
      ms_kernels[3 + j] = b2b[j];
    }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (ll_total) *ll_total = tot;
  return SSM_OK;
}

namespace ssm {
int bf_device_view(ssm_bf* bf, BfView* v) {
  if (!bf || !v) return SSM_ERR_ARG;
  if (!bf->initialised)
    return set_error(bf->ctx, SSM_ERR_STATE, "path storage needs an initialised filter");
  v->ctx = bf->ctx;
  v->N = bf->N;
  v->Npad = bf->Npad;
  v->D = bf->D;
  v->prec = bf->prec;
  v->x = bf->x[bf->cur];
  v->logw = bf->logw;
  v->anc = bf->anc;
  v->sc = bf->sc;
  v->counts_fresh = bf->counts_fresh;
  return SSM_OK;
}

int bf_attach_history(ssm_bf* bf, ssm_history* h, bool on) {
  auto& v = bf->histories;
  v.erase(std::remove(v.begin(), v.end(), h), v.end());
  if (on) v.push_back(h);
  if (!on && bf->recording == h) {
    bf->recording = nullptr;
    destroy_graphs(bf);  // captured pushes write into h
  }
  return SSM_OK;
}

int bf_record_history(ssm_bf* bf, ssm_history* h, bool on) {
  ssm_history* next = on ? h : (bf->recording == h ? nullptr : bf->recording);
  if (next != bf->recording) destroy_graphs(bf);
  bf->recording = next;
  return SSM_OK;
}

int bf_attach_tree(ssm_bf* bf, ssm_tree* tree, int32_t* counts) {
  if (tree && bf->tree && bf->tree != tree)
    return set_error(bf->ctx, SSM_ERR_STATE, "the filter already has path storage attached");
  bf->tree = tree;
  bf->tree_counts = tree ? counts : nullptr;
  bf->counts_fresh = false;
  return SSM_OK;
}
}  // namespace ssm

extern "C" int32_t ssm_bf_last_timing(ssm_bf* bf, double* ms_total, double* ms_per_step) {
  if (!bf) return SSM_ERR_ARG;
  if (ms_total) *ms_total = bf->last_ms;
  if (ms_per_step) *ms_per_step = bf->last_T ? bf->last_ms / bf->last_T : 0.0;
  return SSM_OK;
}

// ---- conditional SMC ----------------------------------------------------------------

namespace {

template <typename T>
int upload_reference_t(ssm_bf* bf, int T_ref, const double* x_ref) {
  ssm_ctx* ctx = bf->ctx;
  const int D = bf->D;
  const size_t n = (size_t)(T_ref + 1) * D;
  if (n > bf->xref_cap) {
    destroy_graphs(bf);  // graphs hold the old buffer address
    if (bf->xref) (void)hipFree(bf->xref);
    bf->xref = nullptr;
    bf->xref_cap = 0;
    if (hipMalloc(&bf->xref, sizeof(T) * n) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "reference trajectory of %d steps", T_ref);
    bf->xref_cap = n;
  }
  std::vector<T> h(n);
  for (size_t e = 0; e < n; ++e) h[e] = (T)x_ref[e];
  if (x_ref != bf->xref_host.data()) bf->xref_host.assign(x_ref, x_ref + n);
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->xref, h.data(), sizeof(T) * n, hipMemcpyHostToDevice,
                                    ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));  // h is a host local
  return upload_transition_t<T>(bf);
}

// RefModelDev: L_Q^{-1} and -1/2 (D log 2 pi + log det Q), from the FP64 Q.
template <typename T>
int upload_transition_t(ssm_bf* bf) {
  ssm_ctx* ctx = bf->ctx;
  const int D = bf->D;
  std::vector<double> LQ = bf->Q_host;
  if (!chol_lower(D, LQ)) return set_error(ctx, SSM_ERR_NOT_PD, "Q is not positive definite");
  const std::vector<double> LQi = tri_inverse(D, LQ);
  double logdet = 0.0;
  for (int q = 0; q < D; ++q) logdet += 2.0 * std::log(LQ[q + q * D]);
  RefModelDev<T> rm;
  std::memset(&rm, 0, sizeof rm);
  for (int e = 0; e < D * D; ++e) rm.LQi[e] = (T)LQi[e];
  rm.cst = (T)(-0.5 * ((double)D * 1.8378770664093454835606594728112 + logdet));
  hipStream_t s = ctx->stream;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->ref_model, &rm, sizeof rm, hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  bf->trans_ok = true;
  return SSM_OK;
}

int ensure_csmc_buffers(ssm_bf* bf) {
  if (bf->ref_model) return SSM_OK;
  ssm_ctx* ctx = bf->ctx;
  const size_t es = (bf->prec == SSM_F32) ? 4 : 8;
  const bool ok =
      hipMalloc(&bf->ref_model, (bf->prec == SSM_F32) ? sizeof(RefModelDev<float>)
                                                      : sizeof(RefModelDev<double>)) ==
          hipSuccess &&
      hipMalloc(&bf->as_w, es * bf->Npad) == hipSuccess &&
      hipMalloc(&bf->as_scratch, resample_scratch_bytes(bf->N)) == hipSuccess &&
      hipMalloc(&bf->as_sc, sizeof(BfScalars)) == hipSuccess &&
      hipMalloc(&bf->as_idx, sizeof(int32_t)) == hipSuccess;
  if (!ok) return set_error(ctx, SSM_ERR_OOM, "conditional SMC buffers");
  BfScalars h;
  std::memset(&h, 0, sizeof h);
  h.ess = (double)bf->N;
  SSM_HIP_CHECK(ctx, hipMemcpy(bf->as_sc, &h, sizeof h, hipMemcpyHostToDevice));
  SSM_HIP_CHECK(ctx, hipMemset(bf->as_idx, 0, sizeof(int32_t)));
  return SSM_OK;
}

}  // namespace

extern "C" int32_t ssm_bf_set_reference(ssm_bf* bf, int32_t T, const double* x_ref,
                                        int32_t mode) {
  if (bf && bf->apf && x_ref)
    return set_error(bf->ctx, SSM_ERR_ARG,
                     "ssm_bf_set_reference: not with the auxiliary particle filter");
  if (!bf) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (!x_ref) {  // back to the unconditional filter
    if (bf->ref_T >= 0) destroy_graphs(bf);
    bf->ref_T = -1;
    return SSM_OK;
  }
  if (T < 0) return set_error(ctx, SSM_ERR_ARG, "reference: T must be >= 0 (got %d)", T);
  if (!bf->prop_F.empty())
    return set_error(ctx, SSM_ERR_ARG, "reference: conditional SMC runs the bootstrap proposal");
  if (mode != SSM_CSMC_PLAIN && mode != SSM_CSMC_ANCESTOR)
    return set_error(ctx, SSM_ERR_ARG, "reference: unknown refreshment %d", mode);
  int rc = ensure_csmc_buffers(bf);
  if (rc) return rc;
  rc = (bf->prec == SSM_F32) ? upload_reference_t<float>(bf, T, x_ref)
                             : upload_reference_t<double>(bf, T, x_ref);
  if (rc) return rc;
  if (bf->ref_T < 0 || bf->ref_mode != mode) destroy_graphs(bf);
  bf->ref_T = T;
  bf->ref_mode = mode;
  return SSM_OK;
}

namespace {
// PropDev<T> from the stored proposal and the current model's Q: L_P, L_Q^-1 and
// -1/2 (log det Q - log det P), in FP64 then rounded to T.
template <typename T>
int upload_proposal_t(ssm_bf* bf) {
  ssm_ctx* ctx = bf->ctx;
  const int D = bf->D, Dy = bf->Dy;
  std::vector<double> LQ = bf->Q_host, LP = bf->prop_P;
  if (!chol_lower(D, LQ)) return set_error(ctx, SSM_ERR_NOT_PD, "Q is not positive definite");
  if (!chol_lower(D, LP))
    return set_error(ctx, SSM_ERR_NOT_PD, "the proposal covariance P is not positive definite");
  const std::vector<double> LQi = tri_inverse(D, LQ);
  double ldq = 0.0, ldp = 0.0;
  for (int q = 0; q < D; ++q) {
    ldq += 2.0 * std::log(LQ[q + q * D]);
    ldp += 2.0 * std::log(LP[q + q * D]);
  }
  if (!bf->prop_dev &&
      hipMalloc(&bf->prop_dev, (bf->prec == SSM_F32) ? sizeof(PropDev<float>)
                                                     : sizeof(PropDev<double>)) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "proposal constants");
  PropDev<T> h;
  std::memset(&h, 0, sizeof h);
  for (int e = 0; e < D * D; ++e) {
    h.F[e] = (T)bf->prop_F[e];
    h.LP[e] = (T)LP[e];
    h.LQi[e] = (T)LQi[e];
  }
  for (int e = 0; e < D; ++e) h.g0[e] = (T)bf->prop_g0[e];
  for (int e = 0; e < D * Dy; ++e) h.K[e] = (T)bf->prop_K[e];
  h.dcst = (T)(-0.5 * (ldq - ldp));
  hipStream_t s = ctx->stream;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->prop_dev, &h, sizeof h, hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));  // h is a host local
  return SSM_OK;
}
}  // namespace

extern "C" int32_t ssm_bf_set_proposal(ssm_bf* bf, const double* F, const double* g0,
                                       const double* K, const double* P) {
  if (bf && bf->apf && F)
    return set_error(bf->ctx, SSM_ERR_ARG,
                     "ssm_bf_set_proposal: not with the auxiliary particle filter");
  if (!bf) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (!F) {  // back to the bootstrap filter
    if (!bf->prop_F.empty()) destroy_graphs(bf);
    bf->prop_F.clear();
    return SSM_OK;
  }
  if (!g0 || !P) return set_error(ctx, SSM_ERR_ARG, "proposal: F, g0 and P are required");
  if (bf->ref_T >= 0)
    return set_error(ctx, SSM_ERR_ARG, "proposal: conditional SMC runs the bootstrap proposal");
  const int D = bf->D, Dy = bf->Dy;
  const bool was = !bf->prop_F.empty();
  std::vector<double> oF = bf->prop_F, og = bf->prop_g0, oK = bf->prop_K, oP = bf->prop_P;
  bf->prop_F.assign(F, F + (size_t)D * D);
  bf->prop_g0.assign(g0, g0 + D);
  if (K)
    bf->prop_K.assign(K, K + (size_t)D * Dy);
  else
    bf->prop_K.assign((size_t)D * Dy, 0.0);
  bf->prop_P.assign(P, P + (size_t)D * D);
  const int rc = (bf->prec == SSM_F32) ? upload_proposal_t<float>(bf) : upload_proposal_t<double>(bf);
  if (rc) {  // keep the previous proposal
    bf->prop_F = oF;
    bf->prop_g0 = og;
    bf->prop_K = oK;
    bf->prop_P = oP;
    return rc;
  }
  if (!was) destroy_graphs(bf);  // the step kernel changes
  return SSM_OK;
}

extern "C" int32_t ssm_bf_set_seed(ssm_bf* bf, uint64_t seed) {
  if (!bf) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // the step and resampler kernels read the key from seed_dev, so captured graphs stay valid
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(bf->seed_dev, &seed, sizeof seed, hipMemcpyHostToDevice,
                                    ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  bf->seed = seed;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_set_auxiliary(ssm_bf* bf, int32_t mode) {
  if (!bf) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  if (mode != SSM_APF_NONE && mode != SSM_APF_MEAN)
    return set_error(ctx, SSM_ERR_ARG, "ssm_bf_set_auxiliary: unknown look-ahead %d", mode);
  if (mode == SSM_APF_MEAN && (bf->ref_T >= 0 || !bf->prop_F.empty()))
    return set_error(ctx, SSM_ERR_ARG,
                     "ssm_bf_set_auxiliary: the auxiliary filter takes neither a reference "
                     "trajectory nor a proposal");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (mode == SSM_APF_MEAN) {
    const int rc = ensure_csmc_buffers(bf);
    if (rc) return rc;
  }
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  destroy_graphs(bf);  // the captured steps differ
  bf->apf = mode == SSM_APF_MEAN;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_sample_index(ssm_bf* bf, double u, int32_t* index) {
  if (!bf || !index) return SSM_ERR_ARG;
  ssm_ctx* ctx = bf->ctx;
  if (!bf->initialised) return set_error(ctx, SSM_ERR_STATE, "ssm_bf_sample_index before initialise");
  if (!(u >= 0.0 && u < 1.0)) return set_error(ctx, SSM_ERR_ARG, "sample_index: u outside [0, 1)");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (!bf->draw_idx && hipMalloc(&bf->draw_idx, sizeof(int32_t)) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "sample_index");
  ResampleArgs a;
  a.kind = SSM_MULTINOMIAL;
  a.n = bf->N;
  a.weight_mode = (bf->prec == SSM_F32) ? kWeightsLogF32 : kWeightsLogF64;
  a.w = bf->logw;
  a.seed = bf->seed;
  a.seed_dev = bf->seed_dev;
  a.sc = bf->sc;
  resample_bind_scratch(a, bf->rs_scratch, bf->N);
  hipStream_t s = ctx->stream;
  int rc = launch_categorical_one(ctx, s, a, u, kStreamAncestor, bf->draw_idx);
  if (rc) return rc;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(index, bf->draw_idx, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return SSM_OK;
}

// ---- backward simulation over a dense history (CSMC-BS) -------------------------------

namespace ssm {

namespace {

// md_tv / rm_tv (time-varying dynamics): Tn entries each, entry t the transition of step
// t + 1 that weights time t (null: the filter's current homogeneous model).
template <typename T>
int backward_sample_t(ssm_bf* bf, const BsHistory& hv, const double* u,
                      const BfModelDev<T>* md_tv = nullptr,
                      const RefModelDev<T>* rm_tv = nullptr) {
  ssm_ctx* ctx = bf->ctx;
  hipStream_t s = ctx->stream;
  const int N = bf->N, Npad = bf->Npad, D = bf->D, Tn = hv.T;
  const size_t plane = (size_t)D * Npad;
  const T* x0 = (const T*)hv.x0;
  const T* st = (const T*)hv.states;
  const int nb = (N + kTile - 1) / kTile;
  const double logN = std::log((double)N);
  ResampleArgs a;
  a.kind = SSM_MULTINOMIAL;
  a.n = N;
  a.weight_mode = (bf->prec == SSM_F32) ? kWeightsLogF32 : kWeightsLogF64;
  a.w = bf->as_w;
  a.seed = bf->seed;
  a.seed_dev = bf->seed_dev;
  a.sc = bf->as_sc;
  resample_bind_scratch(a, bf->as_scratch, N);
  const TileScratch ts{a.partials, a.blk_m, a.blk_c, a.prefix, a.groups};
  for (int t = Tn; t >= 0; --t) {
    const T* x = (t == 0) ? x0 : st + (size_t)(t - 1) * plane;
    const double* lw = (t == 0) ? nullptr : hv.logw + (size_t)(t - 1) * Npad;
    const T* xn = (t == Tn) ? nullptr : st + (size_t)t * plane;  // generation t + 1
    const int32_t* in = bf->bs_idx + t + 1;  // read only when xn is set (t < Tn)
    // the transition x_t -> x_{t+1} (unused at t = Tn)
    const BfModelDev<T>* md =
        (md_tv && t < Tn) ? md_tv + t : (const BfModelDev<T>*)bf->model_dev;
    const RefModelDev<T>* rm =
        (rm_tv && t < Tn) ? rm_tv + t : (const RefModelDev<T>*)bf->ref_model;
    if (D == 4)
      hipLaunchKernelGGL((k_bs_weights<T, 4>), dim3(nb), dim3(kTileThreads), 0, s, N, Npad, D,
                         md, rm, x, lw, xn, in, (T*)bf->as_w, bf->as_sc, a.partials);
    else
      hipLaunchKernelGGL((k_bs_weights<T, 0>), dim3(nb), dim3(kTileThreads), 0, s, N, Npad, D,
                         md, rm, x, lw, xn, in, (T*)bf->as_w, bf->as_sc, a.partials);
    launch_pf_finish<T>(s, N, nb, bf->as_sc, INFINITY, logN, ts, nullptr, nullptr, 0);
    const int rc = launch_categorical_one(ctx, s, a, u[t], kStreamAncestor, bf->bs_idx + t);
    if (rc) return rc;
  }
  const int ne = (Tn + 1) * D;
  hipLaunchKernelGGL(k_bs_gather<T>, dim3((ne + 255) / 256), dim3(256), 0, s, Tn, D, Npad, x0,
                     st, bf->bs_idx, bf->bs_traj);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}

}  // namespace

// The per-step transition constants of time-varying dynamics (dyn[t] the parameters of step
// t + 1, t < T), uploaded into one device block for a backward pass.
template <typename T>
int upload_transitions_tv(ssm_bf* bf, int Tn, const ssm_lg_model* dyn, void** out) {
  ssm_ctx* ctx = bf->ctx;
  const int D = bf->D;
  std::vector<BfModelDev<T>> md(Tn);
  std::vector<RefModelDev<T>> rm(Tn);
  for (int t = 0; t < Tn; ++t) {
    const ssm_lg_model& m = dyn[t];
    if (m.Dx != D || !m.A || !m.b || !m.Q)
      return set_error(ctx, SSM_ERR_ARG, "backward simulation: dynamics %d: bad dimensions", t);
    std::memset(&md[t], 0, sizeof(BfModelDev<T>));
    for (int e = 0; e < D * D; ++e) md[t].A[e] = (T)m.A[e];
    for (int e = 0; e < D; ++e) md[t].b[e] = (T)m.b[e];
    std::vector<double> LQ(m.Q, m.Q + (size_t)D * D);
    if (!chol_lower(D, LQ))
      return set_error(ctx, SSM_ERR_NOT_PD, "backward simulation: Q of step %d is not positive "
                       "definite", t + 1);
    const std::vector<double> LQi = tri_inverse(D, LQ);
    double logdet = 0.0;
    for (int q = 0; q < D; ++q) logdet += 2.0 * std::log(LQ[q + q * D]);
    std::memset(&rm[t], 0, sizeof(RefModelDev<T>));
    for (int e = 0; e < D * D; ++e) rm[t].LQi[e] = (T)LQi[e];
    rm[t].cst = (T)(-0.5 * ((double)D * 1.8378770664093454835606594728112 + logdet));
  }
  const size_t bm = sizeof(BfModelDev<T>) * Tn, br = sizeof(RefModelDev<T>) * Tn;
  if (hipMalloc(out, bm + br) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "backward simulation: %d transitions", Tn);
  hipStream_t s = ctx->stream;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(*out, md.data(), bm, hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync((char*)*out + bm, rm.data(), br, hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));  // md and rm are host locals
  return SSM_OK;
}

int bf_backward_sample(ssm_bf* bf, const BsHistory& hv, const double* u, int32_t* idx,
                       double* traj, const ssm_lg_model* dyn) {
  ssm_ctx* ctx = bf->ctx;
  for (int t = 0; t <= hv.T; ++t)
    if (!(u[t] >= 0.0 && u[t] < 1.0))
      return set_error(ctx, SSM_ERR_ARG, "backward simulation: u[%d] outside [0, 1)", t);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int rc = ensure_csmc_buffers(bf);
  if (rc) return rc;
  if (!bf->trans_ok) {
    rc = (bf->prec == SSM_F32) ? upload_transition_t<float>(bf) : upload_transition_t<double>(bf);
    if (rc) return rc;
  }
  if (hv.T + 1 > bf->bs_cap) {
    if (bf->bs_idx) (void)hipFree(bf->bs_idx);
    if (bf->bs_traj) (void)hipFree(bf->bs_traj);
    bf->bs_idx = nullptr;
    bf->bs_traj = nullptr;
    bf->bs_cap = 0;
    const int cap = std::max(hv.T + 1, 64);
    if (hipMalloc(&bf->bs_idx, sizeof(int32_t) * (cap + 1)) != hipSuccess ||
        hipMalloc(&bf->bs_traj, sizeof(double) * cap * bf->D) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "backward simulation buffers");
    bf->bs_cap = cap;
  }
  void* tv = nullptr;
  if (dyn) {
    rc = (bf->prec == SSM_F32) ? upload_transitions_tv<float>(bf, hv.T, dyn, &tv)
                               : upload_transitions_tv<double>(bf, hv.T, dyn, &tv);
    if (rc) return rc;
  }
  if (bf->prec == SSM_F32) {
    const BfModelDev<float>* md = (const BfModelDev<float>*)tv;
    rc = backward_sample_t<float>(
        bf, hv, u, md, tv ? (const RefModelDev<float>*)(md + hv.T) : nullptr);
  } else {
    const BfModelDev<double>* md = (const BfModelDev<double>*)tv;
    rc = backward_sample_t<double>(
        bf, hv, u, md, tv ? (const RefModelDev<double>*)(md + hv.T) : nullptr);
  }
  hipStream_t s = ctx->stream;
  if (tv) {
    (void)hipStreamSynchronize(s);  // the kernels read tv
    (void)hipFree(tv);
  }
  if (rc) return rc;
  if (idx)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(idx, bf->bs_idx, sizeof(int32_t) * (hv.T + 1),
                                      hipMemcpyDeviceToHost, s));
  if (traj)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(traj, bf->bs_traj, sizeof(double) * (hv.T + 1) * bf->D,
                                      hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return SSM_OK;
}

}  // namespace ssm
