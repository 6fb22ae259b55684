/*
 * ssmhip.h -- C-ABI of the MI355X (gfx950) filtering library.
 *
 * A drop-in beneath GeneralisedFilters' unchanged `filter(model, algo, observations)`:
 * a host (a Julia package extension calling `ccall`, or the Python mirror shipped with
 * this library) keeps the model definitions and the `filter`/`step` driver, reads the
 * model parameters through the unchanged `calc_params` protocol, and hands plain
 * pointers and sizes to these entry points.  No torch or Julia types cross the boundary.
 *
 * Conventions
 *   - Every entry point returns int32_t status: SSM_OK (0) or a negative SSM_ERR_*;
 *     ssm_last_error(ctx) returns the message of the last failure on that context.
 *   - Indices are 0-based int32 (the reference is 1-based Int64; a Julia shim adds 1).
 *   - Matrices are column-major FP64 (Julia/LAPACK layout): M[i + j*rows].
 *   - Host arrays are read (or written) only during the call; device state lives behind
 *     opaque handles until the matching *_destroy.  A context is not thread-safe; use one
 *     per host thread (the serving loop owns its own).
 *   - "dev" pointers are device (HBM) addresses from ssm_dev_alloc or any HIP allocation
 *     on the context's device.
 *
 * Reference paths below are relative to the TuringLang/SSMProblems.jl checkout.
 */
#ifndef SSMHIP_H
#define SSMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (Julia exceptions they map to, in parentheses) ----------------- */
#define SSM_OK 0
#define SSM_ERR_ARG -1        /* ArgumentError / BoundsError                           */
#define SSM_ERR_NOT_PD -2     /* PosDefException from PDMat/cholesky                  */
#define SSM_ERR_OOM -3        /* OutOfMemoryError                                      */
#define SSM_ERR_ADMISSION -4  /* request exceeds the serving HBM budget               */
#define SSM_ERR_HIP -5        /* HIP runtime failure                                   */
#define SSM_ERR_STATE -6      /* call out of order (e.g. step before initialise)      */
#define SSM_ERR_NO_DEVICE -7  /* no HIP device: the library has no CPU fallback       */

/* ---- resampler kinds (GF/src/resamplers.jl:204-240, GF/ext/CUDAExt.jl:31-78) ------ */
#define SSM_SYSTEMATIC 0        /* resamplers.jl:216-238, CPU definition (default)     */
#define SSM_SYSTEMATIC_MURRAY 1 /* CUDAExt.jl:41-58, GPU definition (Murray Code 8)   */
#define SSM_STRATIFIED 2        /* CUDAExt.jl:60-78 (Murray Code 7)                    */
#define SSM_MULTINOMIAL 3       /* CUDAExt.jl:31-39 (inverse CDF, Murray Code 5)       */

/* Conditional SMC trajectory refreshment (GF/src/algorithms/csmc.jl:13-40). */
#define SSM_CSMC_PLAIN 0     /* NoRefreshment: the reference keeps ancestor 0 */
#define SSM_CSMC_ANCESTOR 1  /* AncestorSampling: its ancestor is drawn each step */

/* Auxiliary particle filter look-ahead (GF/src/algorithms/particles.jl:247-393). */
#define SSM_APF_NONE 0 /* the bootstrap filter                                           */
#define SSM_APF_MEAN 1 /* RepresentativeStateLookAhead(MeanPredictive()); for Gaussian     */
                       /* transitions also ModePredictive (the mode is the mean)           */

/* ---- arithmetic types ------------------------------------------------------------- */
#define SSM_F32 0
#define SSM_F64 1

typedef struct ssm_ctx ssm_ctx;
typedef struct ssm_bf ssm_bf;
typedef struct ssm_rbpf ssm_rbpf;
typedef struct ssm_server ssm_server;
typedef struct ssm_kf_batch ssm_kf_batch;
typedef struct ssm_tree ssm_tree;
typedef struct ssm_history ssm_history;

/*
 * Homogeneous linear-Gaussian state-space model, the parameters returned by
 * calc_initial / calc_params (GF/src/models/linear_gaussian.jl:35-63, :96-131).
 * Dx state dim, Dy observation dim; all column-major FP64.
 */
typedef struct {
  int32_t Dx, Dy;
  const double* mu0;    /* Dx      */
  const double* Sigma0; /* Dx x Dx */
  const double* A;      /* Dx x Dx */
  const double* b;      /* Dx      */
  const double* Q;      /* Dx x Dx */
  const double* H;      /* Dy x Dx */
  const double* c;      /* Dy      */
  const double* R;      /* Dy x Dy */
} ssm_lg_model;

/* ---- context ---------------------------------------------------------------------- */
int32_t ssm_version(void);
/* Creates a context on HIP device `device` with its own stream.  hbm_budget_bytes = 0
 * means no budget (the serving loop uses it for admission). */
int32_t ssm_ctx_create(int32_t device, uint64_t hbm_budget_bytes, ssm_ctx** out);
int32_t ssm_ctx_destroy(ssm_ctx* ctx);
const char* ssm_last_error(const ssm_ctx* ctx);
int32_t ssm_ctx_synchronize(ssm_ctx* ctx);
/* Device name and compute-unit count, for reports. */
int32_t ssm_ctx_device_info(ssm_ctx* ctx, char* name, int32_t name_len, int32_t* cu_count);

/* Device buffers for hosts that stage their own data (Julia ROCArray users need none). */
int32_t ssm_dev_alloc(ssm_ctx* ctx, uint64_t bytes, void** out);
int32_t ssm_dev_free(ssm_ctx* ctx, void* p);
int32_t ssm_memcpy_h2d(ssm_ctx* ctx, void* dst_dev, const void* src_host, uint64_t bytes);
int32_t ssm_memcpy_d2h(ssm_ctx* ctx, void* dst_host, const void* src_dev, uint64_t bytes);

/* ---- resampling (GF/src/resamplers.jl:208-238, GF/ext/CUDAExt.jl:31-78) ----------- */
/*
 * sample_ancestors(rng, resampler, weights) for device weights: replaces the CuVector
 * methods at CUDAExt.jl:31-78 and the CPU methods at resamplers.jl:208-238.
 *   d_w      n weights (FP64 or FP32 per w_dtype), need not be normalised for the
 *            Murray/stratified/multinomial forms (they divide by the scan total, as the
 *            reference does); the CPU systematic form assumes sum(w) = 1 as the reference.
 *   d_u      uniforms in [0,1): 1 for systematic, n for stratified/multinomial; NULL draws
 *            them in-kernel from Philox4x32-10 keyed by (seed, step) -- the shim turns the
 *            Julia rng into `seed` by drawing one UInt64.
 *   d_ancestors  n int32 outputs, 0-based.  Systematic/stratified outputs are
 *            nondecreasing.  The FP64 prefix sum is computed on the device.
 */
int32_t ssm_sample_ancestors(ssm_ctx* ctx, int32_t kind, int32_t n, const void* d_w,
                             int32_t w_dtype, const double* d_u, uint64_t seed, uint32_t step,
                             int32_t* d_ancestors);
/*
 * ssm_sample_ancestors without any device allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and results as ssm_sample_ancestors for
 * n <= 1048576 (larger n returns SSM_ERR_ARG); the scratch is a fixed ~10 MiB device array
 * per GPU, held until the process ends and shared by every call on that GPU, so calls that
 * overlap on one GPU (several contexts or host threads at once) race.  It calls no hipMalloc,
 * hipFree or allocating library routine.
 */
int32_t ssm_sample_ancestors_noalloc(ssm_ctx* ctx, int32_t kind, int32_t n, const void* d_w,
                                     int32_t w_dtype, const double* d_u, uint64_t seed,
                                     uint32_t step, int32_t* d_ancestors);
/* Same, with host buffers (copied in and out on the context stream). */
int32_t ssm_sample_ancestors_host(ssm_ctx* ctx, int32_t kind, int32_t n, const double* w,
                                  const double* u, int32_t* ancestors);
/* CUDAExt.jl:88-117 offspring_to_ancestors: cumulative offspring counts -> ancestors. */
int32_t ssm_offspring_to_ancestors(ssm_ctx* ctx, int32_t n, const int32_t* d_offspring,
                                   int32_t* d_ancestors);
/* CUDAExt.jl:119-142 ancestors_to_offspring: ancestors -> per-particle offspring counts. */
int32_t ssm_ancestors_to_offspring(ssm_ctx* ctx, int32_t n, const int32_t* d_ancestors,
                                   int32_t* d_offspring);

/* ---- bootstrap particle filter on an LG model ------------------------------------- */
/*
 * Device-resident BootstrapFilter state (GF/src/algorithms/particles.jl:36-245 with
 * containers.jl:125-173): N particles, state SoA planes x[d][i] in `precision`, log-weights,
 * ancestors and the FP64 scalars of marginalise!.  `seed` keys the in-kernel Philox noise.
 */
int32_t ssm_bf_create(ssm_ctx* ctx, const ssm_lg_model* model, int32_t N, int32_t precision,
                      uint64_t seed, ssm_bf** out);
int32_t ssm_bf_destroy(ssm_bf* bf);
/* calc_params(dyn, t) / calc_params(obs, t) for time-varying models: replaces the
 * parameters used by subsequent steps (particles.jl:117-145 call them per step). */
int32_t ssm_bf_set_model(ssm_bf* bf, const ssm_lg_model* model);
/* initialise(rng, prior, BF) (particles.jl:36-50): x_i = mu0 + chol(Sigma0).L z_i.
 * z: N x Dx host normals (row i = particle i) or NULL for in-kernel Philox noise. */
int32_t ssm_bf_initialise(ssm_bf* bf, const double* z);
/*
 * One `step` (particles.jl:147-162): maybe_resample (ESS rule, resamplers.jl:156-183),
 * predict (particles.jl:52-74), update + marginalise! (particles.jl:76-90,
 * containers.jl:160-173).  t is the reference's 1-based iteration.  u (1 or N host
 * uniforms) and z (N x Dx host normals) override the in-kernel Philox draws (parity mode).
 * Outputs ll_inc and whether this step resampled.
 */
int32_t ssm_bf_step(ssm_bf* bf, int32_t t, const double* y, double threshold,
                    int32_t resampler, const double* u, const double* z, double* ll_inc,
                    int32_t* did_resample);
/*
 * filter(model, BF(N; threshold, resampler), ys) from the current state for steps
 * t0 .. t0+T-1 (GeneralisedFilters.jl:58-84): ys is T x Dy (row t = observation t).
 * All steps run device-side from captured HIP graphs; the resample decision is made on
 * the device.  ll_total receives the summed increments; ll_steps (T, optional) each one.
 */
int32_t ssm_bf_run(ssm_bf* bf, int32_t t0, int32_t T, const double* ys, double threshold,
                   int32_t resampler, double* ll_total, double* ll_steps);
/*
 * Materialises the state for callbacks / checkpoints (callbacks.jl:53-110):
 * x (N x Dx, row i = particle i), normalised logw (N), ancestors of the last step (N),
 * and w (N) the weights the last resample consumed.  Any pointer may be NULL.
 */
int32_t ssm_bf_export(ssm_bf* bf, double* x, double* logw, int32_t* ancestors, double* w);
/* Restores a state (x: N x Dx, logw: N normalised log-weights). */
int32_t ssm_bf_import(ssm_bf* bf, const double* x, const double* logw);
/* Device time of the last ssm_bf_run, from HIP events around it (ms). */
int32_t ssm_bf_last_timing(ssm_bf* bf, double* ms_total, double* ms_per_step);
/*
 * Kernel timing with HIP events on the filter's stream.  Runs the steps of ssm_bf_run one
 * kernel at a time with events around each: ms_kernels[0] resampler, [1] fused
 * predict/update step kernel, [2] finish kernel, each the mean per step (ms; includes the
 * launch gap an event pair adds).  Then, on the final state, launches each of the three
 * kernels `reps` times back to back between two events: ms_kernels[3..5] are the mean
 * durations per launch (ms), comparable to a profiler's per-dispatch duration (reps = 0
 * skips this and writes 0).  ll_total as in ssm_bf_run.  ms_kernels holds 6 doubles.
 */
int32_t ssm_bf_profile(ssm_bf* bf, int32_t t0, int32_t T, const double* ys, double threshold,
                       int32_t resampler, int32_t reps, double* ms_kernels, double* ll_total);

/*
 * AuxiliaryParticleFilter(BF, MeanPredictive()) (GF/src/algorithms/particles.jl:316-344)
 * on this filter: mode SSM_APF_MEAN makes every later step (ssm_bf_step, ssm_bf_run) first
 * score each particle with log eta_i = log N(y_t; H (A x_i + b) + c, R), the observation
 * density at its transition mean (compute_logeta, :282-298), then resample with the
 * AuxiliaryResampler (GF/src/resamplers.jl:92-143): the ESS of softmax(log w + log eta)
 * decides, ancestors are drawn from those weights, a resampled particle starts from
 * -log eta of its ancestor, and the step's increment carries the first-stage correction
 * LSE(log w + log eta) - LSE(log w) of construct_new_state.  Steps that do not resample are
 * the bootstrap filter's.  ssm_bf_export's w is the look-ahead weights the last resample
 * used.  SSM_APF_NONE returns to the bootstrap filter.  Not with a reference trajectory or
 * a proposal (SSM_ERR_ARG); ssm_bf_profile is refused while it is on.
 */
int32_t ssm_bf_set_auxiliary(ssm_bf* bf, int32_t mode);

/* ---- conditional SMC on the bootstrap filter (GF/src/algorithms/csmc.jl) ---------- */
/*
 * Sets (x_ref non-NULL) or clears (x_ref NULL) the reference trajectory of conditional SMC
 * (the ref_state keyword of initialise / step / resample, particles.jl:36-66 and
 * resamplers.jl:51-65).  x_ref is (T + 1) x Dx, row t = the reference state at time t
 * (row 0 the time-0 state, as ReferenceTrajectory indexes from 0).  With a reference,
 * initialise puts particle 0 at x_ref[0] and every step t <= T puts it at x_ref[t]; steps
 * past T are refused.  mode SSM_CSMC_PLAIN (NoRefreshment, csmc.jl:239-263): when a step
 * resamples, particle 0's ancestor is particle 0.  mode SSM_CSMC_ANCESTOR (AncestorSampling,
 * csmc.jl:265-306): every step resamples, and particle 0's ancestor is drawn on the device
 * with probability proportional to w_i N(x_ref[t]; A x_i + b, Q) (ancestor_sampling.jl:
 * 64-130), with a uniform from the filter's Philox key.  Set the model before the
 * reference: Q's factor is taken at this call (and again by ssm_bf_set_model).
 */
int32_t ssm_bf_set_reference(ssm_bf* bf, int32_t T, const double* x_ref, int32_t mode);
/*
 * StatsBase.sample(rng, Weights(get_weights(state))) of _sample_trajectory (csmc.jl:111-117):
 * the 0-based index of the first particle whose cumulative normalised weight reaches u
 * (0 <= u < 1), from the current log-weights (all equal after initialise).
 */
int32_t ssm_bf_sample_index(ssm_bf* bf, double u, int32_t* index);
/* Re-keys the filter's Philox draws (initialise, propagation, resampling, ancestor
 * sampling) for the calls that follow: a new sweep's randomness, as the Julia rng moves on
 * between _csmc_sample calls (csmc.jl:450-462).  Not in the reference's API. */
int32_t ssm_bf_set_seed(ssm_bf* bf, uint64_t seed);
/*
 * Guided particle filter: ParticleFilter(N, proposal) (GF/src/algorithms/particles.jl:92-108)
 * with a linear-Gaussian proposal q(x' | x, y) = N(F x + g0 + K y, P) in place of the latent
 * dynamics.  Each step draws x' from q and adds log f(x' | x) - log q(x' | x, y) to the
 * weight before the observation's density, as propogate does (particles.jl:175-195).  The
 * reference's GFTest proposals are of this form: OptimalProposal (GF/src/GFTest/proposals.jl:
 * 11-31; F = (I - K H) A, g0 = (I - K H) b - K c, the Kalman gain K and P the Joseph-form
 * filtered covariance of Q) and OverdispersedProposal (:33-49; F = A, g0 = b, K = 0,
 * P = k Q).  Column-major F (Dx x Dx), g0 (Dx), K (Dx x Dy, NULL for 0), P (Dx x Dx); the
 * constants follow later ssm_bf_set_model calls' Q.  F = NULL returns to the bootstrap
 * proposal.  Not with a reference trajectory (conditional SMC).
 */
int32_t ssm_bf_set_proposal(ssm_bf* bf, const double* F, const double* g0, const double* K,
                            const double* P);

/* ---- path storage of a bootstrap filter (GF/ext/CUDAExt.jl:146-298) --------------- */
/*
 * Sparse ancestry storage on the device (Jacob, Murray & Rubenthaler 2015): the device
 * ParallelParticleTree with the ParallelAncestorCallback behaviour (CUDAExt.jl:279-298) and
 * the paths of get_ancestry (CUDAExt.jl:232-262; GF/src/containers.jl:592-611).  Nodes
 * without live descendants are pruned and their slots reused; new particle i takes the
 * i-th free slot in index order (insert!, CUDAExt.jl:187-211); the pool doubles when short
 * of N free slots (expand!, CUDAExt.jl:213-227).
 *   ssm_bf_tree_create   roots = the filter's current particles (PostInit); capacity >= N
 *   ssm_bf_tree_insert   appends the particles after the filter's last step, each linked to
 *                        its ancestor (PostUpdate; the identity when the step kept them)
 *   ssm_bf_tree_info     pool size, free slots and depth (number of inserts)
 *   ssm_bf_tree_ancestry every leaf's path: paths[(i (depth+1) + t) Dx + d], t = 0..depth
 *   ssm_bf_tree_path     one leaf's path, path[t Dx + d] (get_ancestry(tree, i, T),
 *                        CUDAExt.jl:248-262)
 *   ssm_bf_tree_export   the pool's parents (-1 at roots and unused slots never written),
 *                        offspring counts (capacity entries each) and the N leaf slots: the
 *                        tree's fields parents/offspring/leaves (CUDAExt.jl:146-150), 0-based
 */
int32_t ssm_bf_tree_create(ssm_bf* bf, int64_t capacity, ssm_tree** out);
int32_t ssm_bf_tree_insert(ssm_tree* tree);
int32_t ssm_bf_tree_info(ssm_tree* tree, int64_t* capacity, int64_t* free_slots,
                         int32_t* depth);
int32_t ssm_bf_tree_ancestry(ssm_tree* tree, double* paths);
int32_t ssm_bf_tree_path(ssm_tree* tree, int32_t i, double* path);
int32_t ssm_bf_tree_export(ssm_tree* tree, int32_t* parents, int32_t* offspring,
                           int32_t* leaves);
int32_t ssm_bf_tree_destroy(ssm_tree* tree);

/* ---- dense particle history (GF/src/containers.jl:328-418) ------------------------- */
/*
 * DenseParticleContainer on the device, filled as DenseAncestorCallback does
 * (GF/src/callbacks.jl:53-80): the time-0 states at creation (PostInit), then after each
 * step (PostUpdate) the states, normalised FP64 log-weights and ancestors (the identity when
 * the step kept its particles).  (4 Dx + 12) B per particle and step in FP32.
 *   ssm_bf_history_create     capacity = generations allocated up front (doubles when full)
 *   ssm_bf_history_push       appends the filter's current generation (push!, :364-374)
 *   ssm_bf_history_reset      empties it and stores the filter's current particles as time
 *                             0, keeping the allocation (a new container per CSMC sweep,
 *                             csmc.jl:140-148, without reallocating)
 *   ssm_bf_history_info       generations pushed and allocated
 *   ssm_bf_history_ancestry   paths of particles i0 .. i0+count-1 (get_ancestry, :406-418):
 *                             paths[(k (T+1) + t) Dx + d], t = 0..T
 *   ssm_bf_history_generation generation t: x[i Dx + d], logw[i], ancestors[i] (0-based);
 *                             t = 0 holds states only (Particle(c, t, i), :396-404)
 */
int32_t ssm_bf_history_create(ssm_bf* bf, int32_t capacity, ssm_history** out);
int32_t ssm_bf_history_push(ssm_history* h);
int32_t ssm_bf_history_reset(ssm_history* h);
/* Recording (on = 1): every later step of the filter, per-step or graph-replayed by
 * ssm_bf_run, appends its generation to h on the device, as DenseAncestorCallback does at
 * PostUpdate inside filter(...; callback) (callbacks.jl:53-80); the history grows as
 * needed.  One history records at a time (turning another on replaces it); on = 0 stops.
 * Do not also call ssm_bf_history_push for the same steps.  ssm_bf_profile's steps are
 * not recorded. */
int32_t ssm_bf_history_record(ssm_history* h, int32_t on);
/*
 * Backward simulation over the stored generations (CSMC-BS, GF/src/algorithms/csmc.jl:
 * 392-446 after its forward pass): index T ~ the last generation's weights, then for
 * t = T-1 .. 0 index t ~ w_t^i N(x*_{t+1}; A x_t^i + b, Q) (ancestor_weight,
 * ancestor_sampling.jl:125-130; the time-0 weights are uniform), each the first particle
 * whose cumulative weight reaches u[t] of the total (StatsBase.sample(Weights(.))).  Uses
 * the filter's current model (homogeneous dynamics).  u: T + 1 uniforms in [0, 1), u[t]
 * for time t; indices (T + 1, 0-based, optional) and trajectory ((T + 1) x Dx, row t the
 * state x*_t, optional) receive the draw.  Every draw runs on the device.
 */
int32_t ssm_bf_history_backward_sample(ssm_history* h, const double* u, int32_t* indices,
                                       double* trajectory);
/*
 * The same backward pass for time-varying dynamics: time t < T is weighted with the
 * transition of step t + 1, as csmc.jl:424 calls ancestor_weight(..., t + 1, ...).
 * dyn holds T models, dyn[t] = calc_params(dyn, t + 1) (only Dx, A, b and Q are read; Dx
 * must be the filter's).  SSM_ERR_NOT_PD names a step whose Q is not positive definite.
 */
int32_t ssm_bf_history_backward_sample_tv(ssm_history* h, const ssm_lg_model* dyn,
                                          const double* u, int32_t* indices,
                                          double* trajectory);
int32_t ssm_bf_history_info(ssm_history* h, int32_t* T, int32_t* capacity);
int32_t ssm_bf_history_ancestry(ssm_history* h, int32_t i0, int32_t count, double* paths);
int32_t ssm_bf_history_generation(ssm_history* h, int32_t t, double* x, double* logw,
                                  int32_t* ancestors);
int32_t ssm_bf_history_destroy(ssm_history* h);

/* ---- Rao-Blackwellised particle filter (GF/src/algorithms/rbpf.jl:8-84) ------------ */
/*
 * Hierarchical model (GF/src/models/hierarchical.jl:4-45) with a scalar outer state and a
 * D-dim linear-Gaussian inner state given the outer path:
 *   outer  x0 ~ N(om0, os0),  x_t = oa x_{t-1} + ob + N(0, oq)
 *   inner  z0 ~ N(mu0, Sigma0), z_t = A z_{t-1} + b + C x_{t-1} + N(0, Q)
 *          (calc_b(dyn; prev_outer) = b + C * prev_outer)
 *   obs    y_t = H z_t + c + N(0, exp(rs * x_t) R0)   (calc_R(obs; new_outer))
 * D <= 32, Dy <= 16; matrices column-major FP64 (C is a D-vector).
 */
typedef struct {
  double om0, os0, oa, ob, oq;
  int32_t D, Dy;
  double rs;
  const double *mu0, *Sigma0, *A, *b, *C, *Q, *H, *c, *R0;
} ssm_rb_model;

/* RBPF(BF(N; threshold, resampler), KalmanFilter()) state: N particles each carrying the
 * outer state and an inner (mu, Sigma) in FP32 on the device; `seed` keys Philox. */
int32_t ssm_rbpf_create(ssm_ctx* ctx, const ssm_rb_model* model, int32_t N, uint64_t seed,
                        ssm_rbpf** out);
int32_t ssm_rbpf_destroy(ssm_rbpf* rb);
/* initialise (rbpf.jl:16-22); z: N outer normals or NULL for Philox. */
int32_t ssm_rbpf_initialise(ssm_rbpf* rb, const double* z);
/* One step (particles.jl:147-162 with rbpf.jl:24-84); u / z override the uniforms and the
 * outer noise (parity mode).  Returns SSM_ERR_NOT_PD if an inner S fails to factorise. */
int32_t ssm_rbpf_step(ssm_rbpf* rb, int32_t t, const double* y, double threshold,
                      int32_t resampler, const double* u, const double* z, double* ll_inc,
                      int32_t* did_resample);
/* Steps t0 .. t0+T-1 with device-side resampling decisions; ys is T x Dy row-major. */
int32_t ssm_rbpf_run(ssm_rbpf* rb, int32_t t0, int32_t T, const double* ys, double threshold,
                     int32_t resampler, double* ll_total, double* ll_steps);
/* State export: xo (N), mu (N x D), Sigma (N x D x D, each column-major), normalised logw
 * (N), ancestors of the last step (N), resample weights of the last step-mode resample. */
int32_t ssm_rbpf_export(ssm_rbpf* rb, double* xo, double* mu, double* Sigma, double* logw,
                        int32_t* ancestors, double* w);
int32_t ssm_rbpf_last_timing(ssm_rbpf* rb, double* ms);

/* ---- serving: continuous batching of whole filter runs ---------------------------- */
/*
 * A request is one filter(model, algo, ys) call (GF/src/GeneralisedFilters.jl:58-84) on a
 * linear-Gaussian model: a BootstrapFilter (particles.jl:36-162; 1 <= Dx, Dy <= 16,
 * 1 <= N <= 262144, FP32 particles, Philox noise keyed by `seed`) or a KalmanFilter
 * (kalman.jl:23-79; Dx <= 64, Dy <= 32).  The reference has no server: these entry points
 * are the batched form of those calls, for hosts that issue many likelihood evaluations
 * (an outer parameter sampler).  The server owns a device arena of hbm_budget_bytes and a
 * worker thread per lane; requests are admitted while slots (max_active bootstrap filters)
 * and arena memory last, keep their state device-resident until they finish, and overlap
 * on the device with everything else admitted.  A bootstrap request's log-likelihood is
 * the same bit for bit alone and in any mix.
 */
#define SSM_REQ_BF 0
#define SSM_REQ_KF 1
/* request states reported by ssm_serve_poll / ssm_serve_wait */
#define SSM_SERVE_QUEUED 0
#define SSM_SERVE_RUNNING 1
#define SSM_SERVE_DONE 2
#define SSM_SERVE_FAILED 3
#define SSM_SERVE_CANCELLED 4

typedef struct {
  int32_t kind;        /* SSM_REQ_BF or SSM_REQ_KF */
  ssm_lg_model model;  /* parameters; copied at submit */
  int32_t N;           /* BF: particles */
  int32_t T;           /* number of observations */
  const double* ys;    /* T x Dy, row t = observation t; copied at submit */
  double threshold;    /* BF: ESS threshold (resample iff threshold * N >= ESS) */
  int32_t resampler;   /* BF: SSM_SYSTEMATIC .. SSM_MULTINOMIAL */
  uint64_t seed;       /* BF: Philox key */
} ssm_request;

/* Errors of create and submit are reported on ctx (ssm_last_error). */
int32_t ssm_serve_create(ssm_ctx* ctx, uint64_t hbm_budget_bytes, int32_t max_active,
                         ssm_server** out);
/* Queues a request; SSM_ERR_ADMISSION if it can never fit the budget, SSM_ERR_NOT_PD if a
 * covariance of a BF model does not factorise.  Thread-safe against the worker threads. */
int32_t ssm_serve_submit(ssm_server* sv, const ssm_request* req, int64_t* id);
/* state: SSM_SERVE_*; ll (log-evidence) and latency_ms (submit to completion) once
 * finished; status: SSM_OK or the request's error code. */
int32_t ssm_serve_poll(ssm_server* sv, int64_t id, int32_t* state, double* ll,
                       double* latency_ms, int32_t* status);
/* Blocks until the request finishes or timeout_ms passes (< 0: no timeout). */
int32_t ssm_serve_wait(ssm_server* sv, int64_t id, double timeout_ms, int32_t* state, double* ll,
                       double* latency_ms, int32_t* status);
/* Cancels a queued request, or a running one at its next batch boundary. */
int32_t ssm_serve_cancel(ssm_server* sv, int64_t id);
/* Forgets a finished request (its result can no longer be polled). */
int32_t ssm_serve_release(ssm_server* sv, int64_t id);
/* out[0..n): completed, failed, cancelled, batched BF steps, Kalman groups, arena bytes in
 * use, peak arena bytes, peak concurrent BF filters, arena capacity, queued requests, BF
 * chunks run, BF chunk graphs captured, ms spent capturing them, ms from each chunk's
 * launch to its retirement (summed). */
int32_t ssm_serve_stats(ssm_server* sv, double* out, int32_t n);
/* Cancels what is queued or running, joins the workers and frees the arena. */
int32_t ssm_serve_destroy(ssm_server* sv);

/* ---- Kalman filters, batched (GF/src/algorithms/kalman.jl:16-79) ------------------- */
/*
 * B independent linear-Gaussian filters of uneven size (1 <= Dx <= 64, 1 <= Dy <= 32),
 * each run for T steps from (mu0, Sigma0): filter(model_f, KalmanFilter(), ys_f) for every
 * f, with the Joseph-form update of kalman_gradient.jl:12-29.  Parameters of filter f are
 * packed column-major FP64 at params + param_off[f] in the order
 *     mu0 (Dx), Sigma0 (Dx*Dx), A (Dx*Dx), b (Dx), Q (Dx*Dx), H (Dy*Dx), c (Dy), R (Dy*Dy),
 * and its observations row by row (T x Dy) at ys + y_off[f].  The data is uploaded once
 * and stays device-resident; each run streams it through persistent kernels.
 */
int32_t ssm_kf_batch_create(ssm_ctx* ctx, int32_t B, const int32_t* Dx, const int32_t* Dy,
                            const double* params, const int64_t* param_off, int32_t T,
                            const double* ys, const int64_t* y_off, ssm_kf_batch** out);
/*
 * Runs all filters.  ll_out (B): total log-likelihood of each filter (NaN where S lost
 * positive definiteness; status_out then holds SSM_ERR_NOT_PD and the call returns it).
 * mu_out / Sigma_out (optional): final filtered means and covariances, packed in filter
 * order (sum of Dx, sum of Dx*Dx entries; each Sigma column-major).
 */
int32_t ssm_kf_batch_run(ssm_kf_batch* kb, double* ll_out, int32_t* status_out, double* mu_out,
                         double* Sigma_out);
/* Device time of the last run from HIP events (ms). */
int32_t ssm_kf_batch_last_timing(ssm_kf_batch* kb, double* ms);
/*
 * Autotuner of the persistent Kalman kernels (no counterpart in the reference): times every
 * candidate workgroup shape (waves per workgroup) of each size class on this batch's own
 * filters for T_tune steps, and keeps the fastest for every later run in the process (the
 * server's Kalman groups included).  nw_out (3): the chosen waves per workgroup of the
 * classes Dx <= 16, <= 32, <= 64; ms_out (9, optional): the time of each candidate
 * (NaN for an empty class or an unused candidate slot).
 */
int32_t ssm_kf_batch_tune(ssm_kf_batch* kb, int32_t T_tune, int32_t* nw_out, double* ms_out);
int32_t ssm_kf_batch_destroy(ssm_kf_batch* kb);
/*
 * kf_loglikelihood(mu0, Sigma0, As, bs, Qs, Hs, cs, Rs, ys, jitter) for B filters of one
 * shape (GF/src/integrations/logdensity.jl:149-161): log p(y_1:T) of each, the Kalman
 * recursion with the Joseph-form update and, when jitter > 0, jitter I added to every
 * filtered covariance (kalman.jl:61-66; 0 is the reference's `nothing`).  Host arrays,
 * column-major FP64 matrices, filter-major:
 *   mu0 (B x Dx), Sigma0 (B x Dx x Dx), ys (B x T x Dy, row t = observation t + 1);
 *   A, Q (P x Dx x Dx), b (P x Dx), H (P x Dy x Dx), c (P x Dy), R (P x Dy x Dy) with
 *   P = B T blocks when time_varying = 1 (block f T + t is step t + 1 of filter f, the
 *   reference's As[t + 1] ...) or P = B when time_varying = 0 (one block per filter for
 *   every step).
 * ll_out (B) receives each log-likelihood (NaN where S lost positive definiteness; its
 * status_out entry, optional, is then SSM_ERR_NOT_PD and the call returns it).
 * 1 <= Dx <= 64, 1 <= Dy <= 32.  The inputs are staged in a device block the context keeps
 * (for filters up to 4 x 4, copied in the caller's layout and transposed on the device to
 * the filter-minor layout the one-filter-per-lane kernel reads).
 */
int32_t ssm_kf_loglikelihood(ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy, int32_t T,
                             int32_t time_varying, const double* mu0, const double* Sigma0,
                             const double* A, const double* b, const double* Q, const double* H,
                             const double* c, const double* R, const double* ys, double jitter,
                             double* ll_out, int32_t* status_out);
/*
 * ssm_kf_loglikelihood with its inputs already in device memory (the same layouts; a
 * ROCArray host passes pointer(x)): no host-to-device copy of the parameters or
 * observations, so a sampler that keeps them on the device pays only the kernels.
 * ll_out and status_out are host arrays, as above.
 */
int32_t ssm_kf_loglikelihood_device(ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy, int32_t T,
                                    int32_t time_varying, const double* d_mu0,
                                    const double* d_Sigma0, const double* d_A,
                                    const double* d_b, const double* d_Q, const double* d_H,
                                    const double* d_c, const double* d_R, const double* d_ys,
                                    double jitter, double* ll_out, int32_t* status_out);
/*
 * The reverse rule of kf_loglikelihood (GF/src/integrations/kalman_rrule.jl:22-107, with the
 * backward recursion of GF/src/algorithms/kalman_gradient.jl:103-263): the log-likelihoods
 * as ssm_kf_loglikelihood and their gradients w.r.t. mu0, Sigma0, A, b, Q, H, c and R, each
 * in its input's layout (P blocks; with time_varying = 0 the gradient of the one shared block
 * sums the steps).  Not w.r.t. ys or jitter, as in the reference.  Filters up to
 * Dx = 64, Dy = 32, in FP64: up to Dx = Dy = 4 one per lane, the forward pass caching 80
 * doubles per step on the device; beyond, one workgroup per filter that keeps the filtered
 * state of every step ((T + 1)(Dx + Dx^2) doubles a filter, filters in chunks of at most
 * 2 GiB of device workspace) and recomputes each step's update in the backward pass.  The
 * log-likelihoods are those of this FP64 forward pass.  Gradients of a filter whose status
 * is SSM_ERR_NOT_PD are undefined.  Any g_* may be NULL.
 */
int32_t ssm_kf_loglikelihood_grad(ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy, int32_t T,
                                  int32_t time_varying, const double* mu0, const double* Sigma0,
                                  const double* A, const double* b, const double* Q,
                                  const double* H, const double* c, const double* R,
                                  const double* ys, double jitter, double* ll_out,
                                  int32_t* status_out, double* g_mu0, double* g_Sigma0,
                                  double* g_A, double* g_b, double* g_Q, double* g_H,
                                  double* g_c, double* g_R);
/*
 * Square-root Kalman filter (GF/src/algorithms/srkf.jl:93-176, SRKalmanFilter(; jitter)):
 * filter(model, SRKF(), ys) for B filters of one shape in the ssm_kf_loglikelihood layout,
 * the covariance carried as its upper Cholesky factor U (Sigma = U' U) through QR
 * factorisations of [U_Q; U A'] (predict) and [[U_R, 0], [U H', U]] (update).  Writes the
 * log-likelihoods, the final filtered means mu_out (B x Dx) and factors U_out (B blocks of
 * Dx x Dx, column-major, upper triangular with a positive diagonal).  jitter is added to
 * the factor's diagonal after each update, as srkf_update does (0: none).  Filters up to
 * Dx = 64, Dy = 32 in FP64: up to Dx = Dy = 4 one per lane (Givens rotations), beyond one
 * workgroup per filter (Householder QR).  Outputs of a filter whose status is
 * SSM_ERR_NOT_PD (Sigma0, Q or R not positive definite) are undefined.  mu_out and U_out
 * may be NULL.
 */
int32_t ssm_srkf_filter(ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy, int32_t T,
                        int32_t time_varying, const double* mu0, const double* Sigma0,
                        const double* A, const double* b, const double* Q, const double* H,
                        const double* c, const double* R, const double* ys, double jitter,
                        double* ll_out, int32_t* status_out, double* mu_out, double* U_out);
/* Device time of the last ssm_kf_loglikelihood's kernel on this context, from HIP events
 * (ms; the host-to-device copies of its inputs are outside). */
int32_t ssm_kf_loglikelihood_last_timing(ssm_ctx* ctx, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* SSMHIP_H */
