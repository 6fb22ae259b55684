/*
 * oracle.c -- FP64 CPU restatement of the GeneralisedFilters filtering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is compared
 * against; it is never linked into, loaded by, or called from the product library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * reference checkout).  Conventions:
 *   - indices are 0-based here; the reference is 1-based (containers.jl:80-84),
 *   - matrices are column-major FP64 (Julia / LAPACK convention), M[i + j*rows],
 *   - particle states are AoS, x[i*D + d], as the reference's Vector{Particle}.
 *
 * Parity status of each part is stated in the header of oracle.py.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_ERR_ARG -1
#define ORC_ERR_NOT_PD -2

/* Resampler kinds -- same numbering as the C-ABI (include/ssmhip.h). */
#define ORC_SYSTEMATIC 0        /* resamplers.jl:214-238 (CPU form)          */
#define ORC_SYSTEMATIC_MURRAY 1 /* CUDAExt.jl:41-58 (GPU form, Murray Code 8) */
#define ORC_STRATIFIED 2        /* CUDAExt.jl:60-78 (Murray Code 7)           */
#define ORC_MULTINOMIAL 3       /* CUDAExt.jl:31-39 (inverse CDF, Code 5)     */

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Counter-based; used by the device kernels for  */
/* in-register noise, restated here so the oracle can reproduce that noise.  */
/* ------------------------------------------------------------------------- */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    philox_round(c, k);
  }
  memcpy(out, c, sizeof c);
}

/* Uniform conversions shared with the device code. */
static inline double u24_open0(uint32_t w) { /* (0,1] */
  return (double)((w >> 8) + 1u) * (1.0 / 16777216.0);
}
static inline double u24_closed0(uint32_t w) { /* [0,1) */
  return (double)(w >> 8) * (1.0 / 16777216.0);
}
static inline double u53_closed0(uint32_t a, uint32_t b) { /* [0,1) */
  const uint64_t v = (((uint64_t)a << 32) | b) >> 11;
  return (double)v * (1.0 / 9007199254740992.0);
}
static inline double u53_open0(uint32_t a, uint32_t b) { /* (0,1] */
  const uint64_t v = (((uint64_t)a << 32) | b) >> 11;
  return (double)(v + 1u) * (1.0 / 9007199254740992.0);
}

/* Philox stream ids (word 2 of the counter). */
#define STREAM_PROPAGATE 0u
#define STREAM_RESAMPLE 1u
#define STREAM_INIT 2u
#define STREAM_OUTER 3u

/*
 * Standard normals for particle i at step t in the device's layout.
 * f32 layout: each Philox call gives 4 words -> 2 Box-Muller pairs -> 4 normals
 *             from 24-bit uniforms; call j covers components 4j..4j+3.
 * f64 layout: each call gives 2 normals from 53-bit uniforms; call j covers 2j..2j+1.
 * Box-Muller: z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = sqrt(-2 ln u1) sin(2 pi u2).
 */
void orc_philox_normals(uint64_t seed, uint32_t stream, uint32_t t, int64_t N, int D,
                        int f64_layout, double* z /* N x D AoS */) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t i = 0; i < N; ++i) {
    if (!f64_layout) {
      for (int j = 0; 4 * j < D; ++j) {
        const uint32_t ctr[4] = {(uint32_t)i, t, stream, (uint32_t)j};
        uint32_t w[4];
        orc_philox4x32_10(ctr, key, w);
        for (int p = 0; p < 2; ++p) {
          const double u1 = u24_open0(w[2 * p]), u2 = u24_closed0(w[2 * p + 1]);
          const double r = sqrt(-2.0 * log(u1));
          const double a = 2.0 * M_PI * u2;
          const int d0 = 4 * j + 2 * p;
          if (d0 < D) z[i * D + d0] = r * cos(a);
          if (d0 + 1 < D) z[i * D + d0 + 1] = r * sin(a);
        }
      }
    } else {
      for (int j = 0; 2 * j < D; ++j) {
        const uint32_t ctr[4] = {(uint32_t)i, t, stream, (uint32_t)j};
        uint32_t w[4];
        orc_philox4x32_10(ctr, key, w);
        const double u1 = u53_open0(w[0], w[1]), u2 = u53_closed0(w[2], w[3]);
        const double r = sqrt(-2.0 * log(u1));
        const double a = 2.0 * M_PI * u2;
        z[i * D + 2 * j] = r * cos(a);
        if (2 * j + 1 < D) z[i * D + 2 * j + 1] = r * sin(a);
      }
    }
  }
}

/* Resampling uniforms u_k in [0,1) for step t (k = 0 is the systematic u0). */
void orc_philox_uniforms(uint64_t seed, uint32_t t, int64_t n, double* u) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t k = 0; k < n; ++k) {
    const uint32_t ctr[4] = {(uint32_t)k, t, STREAM_RESAMPLE, 0u};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    u[k] = u53_closed0(w[0], w[1]);
  }
}

/* ------------------------------------------------------------------------- */
/* Resamplers                                                                 */
/* ------------------------------------------------------------------------- */

/* cumsum(weights): sequential left-to-right FP64 sum, as Julia's cumsum. */
static void cumsum(int64_t n, const double* w, double* W) {
  double s = 0.0;
  for (int64_t j = 0; j < n; ++j) {
    s += w[j];
    W[j] = s;
  }
}

/*
 * CUDAExt.jl:90-117 offspring_to_ancestors: offspring is the cumulative offspring count
 * O (O[n-1] == n); ancestors[O[i-1] .. O[i]-1] = i.  Slots left unfilled when O[n-1] < n
 * (the reference leaves them undefined) are clamped to the last particle.
 */
void orc_offspring_to_ancestors(int64_t n, const int32_t* O, int32_t* anc) {
  int64_t start = 0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t fin = O[i];
    if (fin > n) fin = n;
    for (int64_t j = start; j < fin; ++j) anc[j] = (int32_t)i;
    if (fin > start) start = fin;
  }
  for (int64_t j = start; j < n; ++j) anc[j] = (int32_t)(n - 1);
}

/* CUDAExt.jl:119-142 ancestors_to_offspring: offspring[a_i] += 1 (counts, not cumulative). */
void orc_ancestors_to_offspring(int64_t n, const int32_t* anc, int32_t* off) {
  for (int64_t i = 0; i < n; ++i) off[i] = 0;
  for (int64_t i = 0; i < n; ++i) off[anc[i]] += 1;
}

/*
 * resamplers.jl:216-238, Systematic on CPU vectors:
 *   vs = cumsum(w); vs *= n; u0 ~ U[0,1);
 *   a_k = min{ j : vs_j > u0 + (k-1) }   (strict >, sequential pointer)
 * The reference loop has no j <= n guard (@inbounds); when vs_n <= u0 + (n-1) it reads
 * past the end.  Here the pointer is clamped to the last particle (deliberate, documented).
 */
void orc_systematic(int64_t n, const double* w, double u0, int32_t* anc) {
  double* vs = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
  cumsum(n, w, vs);
  for (int64_t j = 0; j < n; ++j) vs[j] = vs[j] * (double)n;
  int64_t idx = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double u = u0 + (double)i;
    while (idx < n - 1 && vs[idx] <= u) ++idx;
    anc[i] = (int32_t)idx;
  }
  free(vs);
}

/*
 * CUDAExt.jl:49-58 (Murray et al. Code 8):  W = cumsum(w); r = (n*W)/W_n;
 * O_j = min(n, floor(r_j + u0)); then offspring_to_ancestors.
 */
void orc_systematic_murray(int64_t n, const double* w, double u0, int32_t* anc) {
  double* W = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
  int32_t* O = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  cumsum(n, w, W);
  const double Wn = W[n - 1];
  for (int64_t j = 0; j < n; ++j) {
    const double r = ((double)n * W[j]) / Wn;
    double f = floor(r + u0);
    if (f > (double)n) f = (double)n;
    if (f < 0) f = 0;
    O[j] = (int32_t)f;
  }
  orc_offspring_to_ancestors(n, O, anc);
  free(W);
  free(O);
}

/*
 * CUDAExt.jl:68-78 (Murray et al. Code 7): u ~ U[0,1)^n; W = cumsum(w); r = (n*W)/W_n;
 * k_j = min(n, floor(r_j + 1)) (1-based); O_j = min(n, floor(r_j + u[k_j])); o2a.
 */
void orc_stratified(int64_t n, const double* w, const double* u, int32_t* anc) {
  double* W = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
  int32_t* O = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  cumsum(n, w, W);
  const double Wn = W[n - 1];
  for (int64_t j = 0; j < n; ++j) {
    const double r = ((double)n * W[j]) / Wn;
    double kf = floor(r + 1.0);
    if (kf > (double)n) kf = (double)n;
    if (kf < 1) kf = 1;
    const int64_t k = (int64_t)kf; /* 1-based */
    double f = floor(r + u[k - 1]);
    if (f > (double)n) f = (double)n;
    if (f < 0) f = 0;
    O[j] = (int32_t)f;
  }
  orc_offspring_to_ancestors(n, O, anc);
  free(W);
  free(O);
}

/*
 * CUDAExt.jl:31-39 (Murray et al. Code 5): W = cumsum(w); u_k ~ U[0,1) * W_n;
 * a_k = searchsortedfirst(W, u_k) = min{ j : W_j >= u_k }.  Not monotone in k.
 */
void orc_multinomial(int64_t n, const double* w, const double* u, int32_t* anc) {
  double* W = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
  cumsum(n, w, W);
  const double Wn = W[n - 1];
  for (int64_t k = 0; k < n; ++k) {
    const double q = u[k] * Wn;
    int64_t lo = 0, hi = n; /* first j in [lo,hi) with W[j] >= q */
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if (W[mid] < q)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo >= n) lo = n - 1;
    anc[k] = (int32_t)lo;
  }
  free(W);
}

int orc_resample(int kind, int64_t n, const double* w, const double* u, int32_t* anc) {
  if (n < 1) return ORC_ERR_ARG;
  switch (kind) {
    case ORC_SYSTEMATIC: orc_systematic(n, w, u[0], anc); return ORC_OK;
    case ORC_SYSTEMATIC_MURRAY: orc_systematic_murray(n, w, u[0], anc); return ORC_OK;
    case ORC_STRATIFIED: orc_stratified(n, w, u, anc); return ORC_OK;
    case ORC_MULTINOMIAL: orc_multinomial(n, w, u, anc); return ORC_OK;
    default: return ORC_ERR_ARG;
  }
}

/* ------------------------------------------------------------------------- */
/* Small dense linear algebra (column-major FP64)                             */
/* ------------------------------------------------------------------------- */

/* Lower Cholesky in place (lower triangle used; upper zeroed). */
static int chol_lower(int n, double* A) {
  for (int j = 0; j < n; ++j) {
    double d = A[j + j * n];
    for (int k = 0; k < j; ++k) d -= A[j + k * n] * A[j + k * n];
    if (!(d > 0.0)) return ORC_ERR_NOT_PD;
    d = sqrt(d);
    A[j + j * n] = d;
    for (int i = j + 1; i < n; ++i) {
      double s = A[i + j * n];
      for (int k = 0; k < j; ++k) s -= A[i + k * n] * A[j + k * n];
      A[i + j * n] = s / d;
    }
    for (int i = 0; i < j; ++i) A[i + j * n] = 0.0;
  }
  return ORC_OK;
}

/* Solve L v = r in place (forward substitution). */
static void trsv_lower(int n, const double* L, double* v) {
  for (int i = 0; i < n; ++i) {
    double s = v[i];
    for (int k = 0; k < i; ++k) s -= L[i + k * n] * v[k];
    v[i] = s / L[i + i * n];
  }
}

/* Solve L^T v = r in place (back substitution). */
static void trsv_lower_t(int n, const double* L, double* v) {
  for (int i = n - 1; i >= 0; --i) {
    double s = v[i];
    for (int k = i + 1; k < n; ++k) s -= L[k + i * n] * v[k];
    v[i] = s / L[i + i * n];
  }
}

/* C(m x n) = A(m x k) * B(k x n) [+ C if accumulate]; optional transposes. */
static void gemm(int m, int n, int k, const double* A, int ta, const double* B, int tb,
                 double* C, int accumulate) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double s = accumulate ? C[i + j * m] : 0.0;
      for (int p = 0; p < k; ++p) {
        const double a = ta ? A[p + i * k] : A[i + p * m];
        const double b = tb ? B[j + p * n] : B[p + j * k];
        s += a * b;
      }
      C[i + j * m] = s;
    }
}

static const double LOG2PI = 1.8378770664093454835606594728112;

/*
 * logpdf(MvNormal(mu, Sigma), y) = -1/2 (d log 2pi + logdet Sigma + r' Sigma^-1 r),
 * r = y - mu, via the lower Cholesky factor L of Sigma (Distributions.jl restated).
 */
static double mvn_logpdf_chol(int d, const double* L, const double* r_in, double* work) {
  memcpy(work, r_in, sizeof(double) * (size_t)d);
  trsv_lower(d, L, work);
  double q = 0.0, ld = 0.0;
  for (int i = 0; i < d; ++i) {
    q += work[i] * work[i];
    ld += log(L[i + i * d]);
  }
  return -0.5 * ((double)d * LOG2PI + 2.0 * ld + q);
}

/* ------------------------------------------------------------------------- */
/* Linear-Gaussian model                                                      */
/* ------------------------------------------------------------------------- */

/* Parameters of a homogeneous LG SSM (linear_gaussian.jl:96-131), column-major. */
typedef struct {
  int32_t Dx, Dy;
  const double *mu0, *Sigma0, *A, *b, *Q, *H, *c, *R;
} orc_lg_model;

/* logsumexp of a vector: m + log(sum(exp(x - m))); -Inf if every entry is -Inf. */
double orc_logsumexp(int64_t n, const double* x) {
  double m = -INFINITY;
  for (int64_t i = 0; i < n; ++i)
    if (x[i] > m) m = x[i];
  if (m == -INFINITY) return -INFINITY;
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += exp(x[i] - m);
  return m + log(s);
}

/* softmax (LogExpFunctions): exp(x - max) / sum. */
static void softmax(int64_t n, const double* x, double* w) {
  double m = -INFINITY;
  for (int64_t i = 0; i < n; ++i)
    if (x[i] > m) m = x[i];
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    w[i] = exp(x[i] - m);
    s += w[i];
  }
  for (int64_t i = 0; i < n; ++i) w[i] /= s;
}

/*
 * BootstrapFilter initialise: particles.jl:36-50, :110-115, :164-173 ->
 * simulate(prior) -> MvNormal(mu0, Sigma0) sample = mu0 + chol(Sigma0).L * z
 * (linear_gaussian.jl:23-29, :73-76).  z is N x Dx AoS standard normals.
 */
int orc_bf_initialise(const orc_lg_model* m, int64_t N, const double* z, double* x) {
  const int D = m->Dx;
  double* L0 = (double*)malloc(sizeof(double) * (size_t)(D * D));
  memcpy(L0, m->Sigma0, sizeof(double) * (size_t)(D * D));
  if (chol_lower(D, L0)) {
    free(L0);
    return ORC_ERR_NOT_PD;
  }
  for (int64_t i = 0; i < N; ++i)
    for (int r = 0; r < D; ++r) {
      double s = m->mu0[r];
      for (int k = 0; k <= r; ++k) s += L0[r + k * D] * z[i * D + k];
      x[i * D + r] = s;
    }
  free(L0);
  return ORC_OK;
}

/*
 * One BootstrapFilter step, particles.jl:147-162 (step) with
 *   maybe_resample  resamplers.jl:156-183 (ESSResampler: resample iff threshold*N >= ESS,
 *                   ESS = 1/sum(w^2), w = softmax(logw); TypelessZero weights -> 1/N,
 *                   containers.jl:62-65), construct_new_state :69-79 (gather, logw = 0),
 *                   preserve_sample :149-154 (a_i = i, baseline kept);
 *   predict         particles.jl:52-74 (x' = A x + b + chol(Q).L z, BF increment 0
 *                   :228-245; baseline = logsumexp(logw) + baseline, TypelessBaseline
 *                   = log N for unweighted particles, containers.jl:43-60);
 *   update          particles.jl:76-90, :133-145 (logw += logpdf(MvNormal(Hx + c, R), y))
 *                   and marginalise! containers.jl:160-173 (ll = LSE - baseline,
 *                   logw -= LSE).
 * State: x (N x Dx AoS) and logw are updated in place; *typeless is 1 while the
 * weights are still the lazy TypelessZero of initialise (particles.jl:49).
 * u: the resampler uniforms (1 for systematic, N otherwise).  z: N x Dx propagate
 * noise.  w_override (optional): weights to hand to the resampler instead of the
 * softmax computed here (teacher forcing with a device's own weights).
 * Outputs: anc (N), *ll_inc, *did_resample, w_out (optional, the weights used).
 */
static int bf_step_core(const orc_lg_model* m, int64_t N, double* x, double* logw,
                        int* typeless, double threshold, int kind, const double* u,
                        const double* z, const double* y, const double* w_override,
                        const double* x_ref_t, int32_t ref_anc, int32_t* anc, double* ll_inc,
                        int* did_resample, double* w_out) {
  const int D = m->Dx, Dy = m->Dy;
  double* w = (double*)malloc(sizeof(double) * (size_t)N);
  double* xn = (double*)malloc(sizeof(double) * (size_t)(N * D));
  double* LQ = (double*)malloc(sizeof(double) * (size_t)(D * D));
  double* LR = (double*)malloc(sizeof(double) * (size_t)(Dy * Dy));
  double* r = (double*)malloc(sizeof(double) * (size_t)Dy);
  double* work = (double*)malloc(sizeof(double) * (size_t)Dy);
  int rc = ORC_OK;
  memcpy(LQ, m->Q, sizeof(double) * (size_t)(D * D));
  memcpy(LR, m->R, sizeof(double) * (size_t)(Dy * Dy));
  if (chol_lower(D, LQ) || chol_lower(Dy, LR)) {
    rc = ORC_ERR_NOT_PD;
    goto done;
  }

  /* maybe_resample */
  if (*typeless) {
    for (int64_t i = 0; i < N; ++i) w[i] = 1.0 / (double)N;
  } else {
    softmax(N, logw, w);
  }
  {
    double s2 = 0.0;
    for (int64_t i = 0; i < N; ++i) s2 += w[i] * w[i];
    const double ess = 1.0 / s2;
    *did_resample = (threshold * (double)N >= ess) ? 1 : 0;
  }
  if (w_out) memcpy(w_out, w, sizeof(double) * (size_t)N);
  double baseline;
  if (*did_resample) {
    rc = orc_resample(kind, N, w_override ? w_override : w, u, anc);
    if (rc) goto done;
    if (x_ref_t) anc[0] = ref_anc; /* _set_ref_index! / the drawn ancestor, resamplers.jl:51-65 */
    for (int64_t i = 0; i < N; ++i) {
      const int64_t a = anc[i];
      for (int d = 0; d < D; ++d) xn[i * D + d] = x[a * D + d];
    }
    memcpy(x, xn, sizeof(double) * (size_t)(N * D));
    for (int64_t i = 0; i < N; ++i) logw[i] = 0.0;
    *typeless = 0;
    baseline = orc_logsumexp(N, logw); /* + ll_baseline (= 0 after construct_new_state) */
  } else {
    for (int64_t i = 0; i < N; ++i) anc[i] = (int32_t)i;
    if (*typeless) {
      baseline = log((double)N); /* TypelessBaseline(N) */
      for (int64_t i = 0; i < N; ++i) logw[i] = 0.0;
      *typeless = 0;
    } else {
      baseline = orc_logsumexp(N, logw); /* + ll_baseline (= 0 after marginalise!) */
    }
  }

  /* predict: x' = A x + b + L_Q z */
  for (int64_t i = 0; i < N; ++i)
    for (int rr = 0; rr < D; ++rr) {
      double s = m->b[rr];
      for (int k = 0; k < D; ++k) s += m->A[rr + k * D] * x[i * D + k];
      double e = 0.0;
      for (int k = 0; k <= rr; ++k) e += LQ[rr + k * D] * z[i * D + k];
      xn[i * D + rr] = s + e;
    }
  memcpy(x, xn, sizeof(double) * (size_t)(N * D));
  /* conditional SMC: particle 1 (index 0) moves to the reference state (particles.jl:59-66) */
  if (x_ref_t)
    for (int d = 0; d < D; ++d) x[d] = x_ref_t[d];

  /* update: logw += logpdf(MvNormal(H x + c, R), y) */
  for (int64_t i = 0; i < N; ++i) {
    for (int q = 0; q < Dy; ++q) {
      double s = m->c[q];
      for (int k = 0; k < D; ++k) s += m->H[q + k * Dy] * x[i * D + k];
      r[q] = y[q] - s;
    }
    logw[i] += mvn_logpdf_chol(Dy, LR, r, work);
  }
  /* marginalise! */
  {
    const double lse = orc_logsumexp(N, logw);
    *ll_inc = lse - baseline;
    for (int64_t i = 0; i < N; ++i) logw[i] -= lse;
  }
done:
  free(w);
  free(xn);
  free(LQ);
  free(LR);
  free(r);
  free(work);
  return rc;
}

int orc_bf_step(const orc_lg_model* m, int64_t N, double* x, double* logw, int* typeless,
                double threshold, int kind, const double* u, const double* z,
                const double* y, const double* w_override, int32_t* anc, double* ll_inc,
                int* did_resample, double* w_out) {
  return bf_step_core(m, N, x, logw, typeless, threshold, kind, u, z, y, w_override, NULL, 0,
                      anc, ll_inc, did_resample, w_out);
}

/*
 * One auxiliary particle filter step with the transition mean as representative state
 * (AuxiliaryParticleFilter with MeanPredictive, GF/src/algorithms/particles.jl:248-305 and
 * step at :316-344; AuxiliaryResampler, GF/src/resamplers.jl:92-143):
 *   log eta_i = logpdf(MvNormal(H (A x_i + b) + c, R), y)      (compute_logeta)
 *   w~ = softmax(logw + log eta); ESS = 1 / sum w~^2; resample iff threshold N >= ESS
 *   resample: a ~ w~, x_i <- x_{a_i}, logw_i <- -log eta_{a_i} (resample_ancestor), and
 *     ll_baseline = -((LSE(log eta + logw) - LSE(logw)) + (LSE(-log eta[a]) - log N))
 *     (construct_new_state, LSE(logw) of the lazy TypelessZero weights being log N);
 *   otherwise preserve_sample: particles and baseline kept.
 * Then predict adds LSE(logw) to the baseline, update adds the observation log-density and
 * marginalise! returns LSE(logw) - baseline (particles.jl:52-90, containers.jl:160-173).
 * u, z, w_override, w_out (the auxiliary weights) as in orc_bf_step.
 */
int orc_apf_step(const orc_lg_model* m, int64_t N, double* x, double* logw, int* typeless,
                 double threshold, int kind, const double* u, const double* z, const double* y,
                 const double* w_override, int32_t* anc, double* ll_inc, int* did_resample,
                 double* w_out) {
  const int D = m->Dx, Dy = m->Dy;
  double* eta = (double*)malloc(sizeof(double) * (size_t)N);
  double* aux = (double*)malloc(sizeof(double) * (size_t)N);
  double* w = (double*)malloc(sizeof(double) * (size_t)N);
  double* xn = (double*)malloc(sizeof(double) * (size_t)(N * D));
  double* LQ = (double*)malloc(sizeof(double) * (size_t)(D * D));
  double* LR = (double*)malloc(sizeof(double) * (size_t)(Dy * Dy));
  double* xm = (double*)malloc(sizeof(double) * (size_t)D);
  double* r = (double*)malloc(sizeof(double) * (size_t)Dy);
  double* work = (double*)malloc(sizeof(double) * (size_t)Dy);
  int rc = ORC_OK;
  memcpy(LQ, m->Q, sizeof(double) * (size_t)(D * D));
  memcpy(LR, m->R, sizeof(double) * (size_t)(Dy * Dy));
  if (chol_lower(D, LQ) || chol_lower(Dy, LR)) {
    rc = ORC_ERR_NOT_PD;
    goto done;
  }
  /* look-ahead scores at the transition means */
  for (int64_t i = 0; i < N; ++i) {
    for (int rr = 0; rr < D; ++rr) {
      double s = m->b[rr];
      for (int k = 0; k < D; ++k) s += m->A[rr + k * D] * x[i * D + k];
      xm[rr] = s;
    }
    for (int q = 0; q < Dy; ++q) {
      double s = m->c[q];
      for (int k = 0; k < D; ++k) s += m->H[q + k * Dy] * xm[k];
      r[q] = y[q] - s;
    }
    eta[i] = mvn_logpdf_chol(Dy, LR, r, work);
    aux[i] = (*typeless ? 0.0 : logw[i]) + eta[i];
  }
  softmax(N, aux, w);
  {
    double s2 = 0.0;
    for (int64_t i = 0; i < N; ++i) s2 += w[i] * w[i];
    *did_resample = (threshold * (double)N >= 1.0 / s2) ? 1 : 0;
  }
  if (w_out) memcpy(w_out, w, sizeof(double) * (size_t)N);
  double baseline;
  if (*did_resample) {
    const double lse1 = orc_logsumexp(N, aux);
    const double lse2 = *typeless ? log((double)N) : orc_logsumexp(N, logw);
    rc = orc_resample(kind, N, w_override ? w_override : w, u, anc);
    if (rc) goto done;
    for (int64_t i = 0; i < N; ++i) {
      const int64_t a = anc[i];
      for (int d = 0; d < D; ++d) xn[i * D + d] = x[a * D + d];
      logw[i] = -eta[a];
    }
    memcpy(x, xn, sizeof(double) * (size_t)(N * D));
    *typeless = 0;
    const double lse3 = orc_logsumexp(N, logw);
    const double beta = -((lse1 - lse2) + (lse3 - log((double)N)));
    baseline = lse3 + beta; /* predict: LSE(logw) + ll_baseline */
  } else {
    for (int64_t i = 0; i < N; ++i) anc[i] = (int32_t)i;
    if (*typeless) {
      baseline = log((double)N);
      for (int64_t i = 0; i < N; ++i) logw[i] = 0.0;
      *typeless = 0;
    } else {
      baseline = orc_logsumexp(N, logw);
    }
  }
  /* predict and update as the bootstrap filter */
  for (int64_t i = 0; i < N; ++i)
    for (int rr = 0; rr < D; ++rr) {
      double s = m->b[rr];
      for (int k = 0; k < D; ++k) s += m->A[rr + k * D] * x[i * D + k];
      double e = 0.0;
      for (int k = 0; k <= rr; ++k) e += LQ[rr + k * D] * z[i * D + k];
      xn[i * D + rr] = s + e;
    }
  memcpy(x, xn, sizeof(double) * (size_t)(N * D));
  for (int64_t i = 0; i < N; ++i) {
    for (int q = 0; q < Dy; ++q) {
      double s = m->c[q];
      for (int k = 0; k < D; ++k) s += m->H[q + k * Dy] * x[i * D + k];
      r[q] = y[q] - s;
    }
    logw[i] += mvn_logpdf_chol(Dy, LR, r, work);
  }
  {
    const double lse = orc_logsumexp(N, logw);
    *ll_inc = lse - baseline;
    for (int64_t i = 0; i < N; ++i) logw[i] -= lse;
  }
done:
  free(eta);
  free(aux);
  free(w);
  free(xn);
  free(LQ);
  free(LR);
  free(xm);
  free(r);
  free(work);
  return rc;
}

/*
 * One conditional-SMC step (csmc.jl:239-306): the BootstrapFilter step above with the
 * reference pinned to particle 1 (index 0).  When the step resamples, particle 0's
 * ancestor is ref_anc (0 for NoRefreshment's _set_ref_index!, resamplers.jl:51-65, or
 * the index ancestor sampling drew, csmc.jl:276-297); after the predict particle 0 is
 * x_ref_t, the reference state at this step (particles.jl:59-66).  Ancestor sampling
 * always resamples (csmc.jl:288 calls resample, not maybe_resample): pass threshold
 * +Inf for it.  w_override and w_out as in orc_bf_step.
 */
int orc_csmc_step(const orc_lg_model* m, int64_t N, double* x, double* logw, int* typeless,
                  double threshold, int kind, const double* u, const double* z,
                  const double* y, const double* w_override, const double* x_ref_t,
                  int32_t ref_anc, int32_t* anc, double* ll_inc, int* did_resample,
                  double* w_out) {
  if (!x_ref_t || ref_anc < 0 || ref_anc >= N) return ORC_ERR_ARG;
  return bf_step_core(m, N, x, logw, typeless, threshold, kind, u, z, y, w_override, x_ref_t,
                      ref_anc, anc, ll_inc, did_resample, w_out);
}

/*
 * Ancestor-sampling log-weights (ancestor_weight, ancestor_sampling.jl:125-130, with the
 * generic future_conditional_density :64-68 = SSMProblems.logdensity of the LG dynamics):
 * out_i = log w_i + log N(x_ref; A x_i + b, Q), log w_i the current log-weights (0 for the
 * TypelessZero weights after initialise).  x is N x Dx AoS.
 */
int orc_as_logweights(const orc_lg_model* m, int64_t N, const double* x, const double* logw,
                      int typeless, const double* x_ref, double* out) {
  const int D = m->Dx;
  double* LQ = (double*)malloc(sizeof(double) * (size_t)(D * D));
  double* r = (double*)malloc(sizeof(double) * (size_t)D);
  double* work = (double*)malloc(sizeof(double) * (size_t)D);
  int rc = ORC_OK;
  memcpy(LQ, m->Q, sizeof(double) * (size_t)(D * D));
  if (chol_lower(D, LQ)) {
    rc = ORC_ERR_NOT_PD;
  } else {
    for (int64_t i = 0; i < N; ++i) {
      for (int a = 0; a < D; ++a) {
        double s = m->b[a];
        for (int k = 0; k < D; ++k) s += m->A[a + k * D] * x[i * D + k];
        r[a] = x_ref[a] - s;
      }
      out[i] = (typeless ? 0.0 : logw[i]) + mvn_logpdf_chol(D, LQ, r, work);
    }
  }
  free(LQ);
  free(r);
  free(work);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* Kalman filter                                                              */
/* ------------------------------------------------------------------------- */

/*
 * kalman_predict, kalman.jl:42-49: mu' = A mu + b, Sigma' = A Sigma A' + Q.
 * (PDMats' X_A_Xt evaluates A Sigma A' through Sigma's Cholesky factor; the plain
 * triple product here is the same quantity up to rounding.)
 */
void orc_kf_predict(int Dx, const double* A, const double* b, const double* Q, double* mu,
                    double* Sigma) {
  double* tmp = (double*)malloc(sizeof(double) * (size_t)(Dx * Dx));
  double* mu2 = (double*)malloc(sizeof(double) * (size_t)Dx);
  for (int i = 0; i < Dx; ++i) {
    double s = b[i];
    for (int k = 0; k < Dx; ++k) s += A[i + k * Dx] * mu[k];
    mu2[i] = s;
  }
  memcpy(mu, mu2, sizeof(double) * (size_t)Dx);
  gemm(Dx, Dx, Dx, A, 0, Sigma, 0, tmp, 0);  /* A Sigma */
  memcpy(Sigma, Q, sizeof(double) * (size_t)(Dx * Dx));
  gemm(Dx, Dx, Dx, tmp, 0, A, 1, Sigma, 1);  /* + (A Sigma) A' */
  free(tmp);
  free(mu2);
}

/*
 * _kalman_update_cached, kalman_gradient.jl:12-23 with kalman.jl:51-66:
 *   z = y - H mu - c;  S = H Sigma H' + R  (PDMat: Cholesky);  K = Sigma H' S^-1;
 *   Joseph form Sigma' = (I - K H) Sigma (I - K H)' + K R K' (+ jitter I), symmetrised;
 *   mu' = mu + K z;  ll = logpdf(MvNormal(z, S), 0).
 * Returns ORC_ERR_NOT_PD if S is not positive definite.
 */
int orc_kf_update(int Dx, int Dy, const double* H, const double* c, const double* R,
                  const double* y, double jitter, double* mu, double* Sigma, double* ll) {
  const size_t nxx = (size_t)(Dx * Dx), nxy = (size_t)(Dx * Dy), nyy = (size_t)(Dy * Dy);
  double* zv = (double*)malloc(sizeof(double) * (size_t)Dy);
  double* S = (double*)malloc(sizeof(double) * nyy);
  double* HS = (double*)calloc(nxy, sizeof(double));     /* H Sigma  (Dy x Dx) */
  double* K = (double*)malloc(sizeof(double) * nxy);    /* Dx x Dy */
  double* IKH = (double*)malloc(sizeof(double) * nxx);
  double* T1 = (double*)malloc(sizeof(double) * nxx);
  double* KR = (double*)malloc(sizeof(double) * nxy);
  double* work = (double*)malloc(sizeof(double) * (size_t)(Dy > Dx ? Dy : Dx));
  int rc = ORC_OK;
  for (int q = 0; q < Dy; ++q) {
    double s = c[q];
    for (int k = 0; k < Dx; ++k) s += H[q + k * Dy] * mu[k];
    zv[q] = y[q] - s;
  }
  gemm(Dy, Dx, Dx, H, 0, Sigma, 0, HS, 0);
  memcpy(S, R, sizeof(double) * nyy);
  gemm(Dy, Dy, Dx, HS, 0, H, 1, S, 1);
  /* X_A_Xt(Sigma, H) is formed as (H L)(H L)' in PDMats, exactly symmetric; only the
   * lower triangle of S is read by the Cholesky factorisation below. */
  double* L = (double*)malloc(sizeof(double) * nyy);
  memcpy(L, S, sizeof(double) * nyy);
  if (chol_lower(Dy, L)) {
    rc = ORC_ERR_NOT_PD;
    goto done;
  }
  /* K' = S^-1 (H Sigma) since Sigma symmetric: solve column by column of HS */
  for (int j = 0; j < Dx; ++j) {
    for (int q = 0; q < Dy; ++q) work[q] = HS[q + j * Dy];
    trsv_lower(Dy, L, work);
    trsv_lower_t(Dy, L, work);
    for (int q = 0; q < Dy; ++q) K[j + q * Dx] = work[q];
  }
  *ll = mvn_logpdf_chol(Dy, L, zv, work);
  /* mu' = mu + K z */
  for (int i = 0; i < Dx; ++i) {
    double s = 0.0;
    for (int q = 0; q < Dy; ++q) s += K[i + q * Dx] * zv[q];
    mu[i] += s;
  }
  /* Joseph form */
  gemm(Dx, Dx, Dy, K, 0, H, 0, IKH, 0);
  for (size_t e = 0; e < nxx; ++e) IKH[e] = -IKH[e];
  for (int i = 0; i < Dx; ++i) IKH[i + i * Dx] += 1.0;
  gemm(Dx, Dx, Dx, IKH, 0, Sigma, 0, T1, 0);       /* (I-KH) Sigma */
  gemm(Dx, Dy, Dy, K, 0, R, 0, KR, 0);             /* K R */
  gemm(Dx, Dx, Dx, T1, 0, IKH, 1, Sigma, 0);       /* (I-KH) Sigma (I-KH)' */
  gemm(Dx, Dx, Dy, KR, 0, K, 1, Sigma, 1);         /* + K R K' */
  if (jitter != 0.0)
    for (int i = 0; i < Dx; ++i) Sigma[i + i * Dx] += jitter;
  /* PDMat(Symmetric(Sigma_raw)) (kalman.jl:65): Symmetric keeps the upper triangle */
  for (int j = 0; j < Dx; ++j)
    for (int i = j + 1; i < Dx; ++i) Sigma[i + j * Dx] = Sigma[j + i * Dx];
done:
  free(zv);
  free(S);
  free(HS);
  free(K);
  free(IKH);
  free(T1);
  free(KR);
  free(work);
  free(L);
  return rc;
}

/*
 * filter(model, KalmanFilter(), ys), GeneralisedFilters.jl:58-84 with kalman.jl:23-79:
 * initialise to (mu0, Sigma0), then T x (predict, update), summing ll.
 * ys is T x Dy (row t = observation t).  mu_out/Sigma_out receive the final filtered state.
 */
int orc_kf_filter(const orc_lg_model* m, int T, const double* ys, double jitter,
                  double* mu_out, double* Sigma_out, double* ll_total, double* ll_steps) {
  const int Dx = m->Dx, Dy = m->Dy;
  memcpy(mu_out, m->mu0, sizeof(double) * (size_t)Dx);
  memcpy(Sigma_out, m->Sigma0, sizeof(double) * (size_t)(Dx * Dx));
  double tot = 0.0;
  for (int t = 0; t < T; ++t) {
    double ll;
    orc_kf_predict(Dx, m->A, m->b, m->Q, mu_out, Sigma_out);
    const int rc = orc_kf_update(Dx, Dy, m->H, m->c, m->R, ys + (size_t)t * Dy, jitter,
                                 mu_out, Sigma_out, &ll);
    if (rc) return rc;
    if (ll_steps) ll_steps[t] = ll;
    tot += ll;
  }
  *ll_total = tot;
  return ORC_OK;
}

/*
 * Batched form: B independent filters of uneven size, parameters packed back to back
 * (per filter: mu0 Dx, Sigma0 DxDx, A DxDx, b Dx, Q DxDx, H DyDx, c Dy, R DyDy), ys
 * packed per filter as T x Dy.  Writes the total log-likelihood of each filter.
 */
int orc_kf_batch(int B, const int32_t* Dx, const int32_t* Dy, const double* params,
                 const int64_t* param_off, const double* ys, const int64_t* y_off, int T,
                 double* ll_out) {
  int rc = ORC_OK;
  for (int f = 0; f < B && rc == ORC_OK; ++f) {
    const int dx = Dx[f], dy = Dy[f];
    const double* p = params + param_off[f];
    orc_lg_model m;
    m.Dx = dx;
    m.Dy = dy;
    m.mu0 = p;
    p += dx;
    m.Sigma0 = p;
    p += dx * dx;
    m.A = p;
    p += dx * dx;
    m.b = p;
    p += dx;
    m.Q = p;
    p += dx * dx;
    m.H = p;
    p += dy * dx;
    m.c = p;
    p += dy;
    m.R = p;
    double* mu = (double*)malloc(sizeof(double) * (size_t)dx);
    double* Sig = (double*)malloc(sizeof(double) * (size_t)(dx * dx));
    rc = orc_kf_filter(&m, T, ys + y_off[f], 0.0, mu, Sig, &ll_out[f], NULL);
    free(mu);
    free(Sig);
  }
  return rc;
}

/* ------------------------------------------------------------------------- */
/* Rao-Blackwellised particle filter (outer scalar, inner linear Gaussian)    */
/* ------------------------------------------------------------------------- */

/*
 * Hierarchical model (hierarchical.jl:4-45) with a 1-D outer state and a D-dim inner
 * LG model conditioned on it:
 *   outer:  x0 ~ N(om0, os0);  x_t = oa x_{t-1} + ob + N(0, oq)
 *   inner:  z0 ~ N(mu0, Sigma0);  z_t = A z_{t-1} + b + C x_{t-1} + N(0, Q)
 *           (calc_b(dyn; prev_outer) = b + C * prev_outer, as the dummy LG model's
 *           InnerDynamics, GFTest/models/dummy_linear_gaussian.jl:43-45)
 *   obs:    y_t = H z_t + c + N(0, exp(rs * x_t) R0)
 *           (rs = 1 gives the exp(outer)-scaled noise of the trend-inflation example,
 *           GF/examples/trend-inflation/script.jl:132-134; rs = 0 gives plain R0).
 */
typedef struct {
  double om0, os0, oa, ob, oq;
  int32_t D, Dy;
  double rs;
  const double *mu0, *Sigma0, *A, *b, *C, *Q, *H, *c, *R0;
} orc_rb_model;

/* RBPF initialise_particle, rbpf.jl:16-22: x ~ outer prior, z = (mu0, Sigma0). */
void orc_rbpf_initialise(const orc_rb_model* m, int64_t N, const double* zo, double* xo,
                         double* mu, double* Sigma) {
  const int D = m->D;
  for (int64_t i = 0; i < N; ++i) {
    xo[i] = m->om0 + sqrt(m->os0) * zo[i];
    memcpy(mu + i * D, m->mu0, sizeof(double) * (size_t)D);
    memcpy(Sigma + i * D * D, m->Sigma0, sizeof(double) * (size_t)(D * D));
  }
}

/*
 * One RBPF step: particles.jl:147-162 with rbpf.jl:24-84.  maybe_resample as the BF;
 * predict_particle: outer BF propagate x' = oa x + ob + sqrt(oq) zo, inner kalman_predict
 * with prev_outer = x, new_outer = x'; update_particle: inner kalman_update with
 * new_outer = x' (R = exp(rs x') R0), logw += inner ll; then marginalise!.
 */
int orc_rbpf_step(const orc_rb_model* m, int64_t N, double* xo, double* mu, double* Sigma,
                  double* logw, int* typeless, double threshold, int kind, const double* u,
                  const double* zo, const double* y, const double* w_override,
                  int32_t* anc, double* ll_inc, int* did_resample, double* w_out) {
  const int D = m->D, Dy = m->Dy;
  double* w = (double*)malloc(sizeof(double) * (size_t)N);
  double* xo2 = (double*)malloc(sizeof(double) * (size_t)N);
  double* mu2 = (double*)malloc(sizeof(double) * (size_t)(N * D));
  double* S2 = (double*)malloc(sizeof(double) * (size_t)(N * D * D));
  double* bt = (double*)malloc(sizeof(double) * (size_t)D);
  double* Rt = (double*)malloc(sizeof(double) * (size_t)(Dy * Dy));
  int rc = ORC_OK;
  if (*typeless) {
    for (int64_t i = 0; i < N; ++i) w[i] = 1.0 / (double)N;
  } else {
    softmax(N, logw, w);
  }
  {
    double s2 = 0.0;
    for (int64_t i = 0; i < N; ++i) s2 += w[i] * w[i];
    *did_resample = (threshold * (double)N >= 1.0 / s2) ? 1 : 0;
  }
  if (w_out) memcpy(w_out, w, sizeof(double) * (size_t)N);
  double baseline;
  if (*did_resample) {
    rc = orc_resample(kind, N, w_override ? w_override : w, u, anc);
    if (rc) goto done;
    for (int64_t i = 0; i < N; ++i) {
      const int64_t a = anc[i];
      xo2[i] = xo[a];
      memcpy(mu2 + i * D, mu + a * D, sizeof(double) * (size_t)D);
      memcpy(S2 + i * D * D, Sigma + a * D * D, sizeof(double) * (size_t)(D * D));
    }
    memcpy(xo, xo2, sizeof(double) * (size_t)N);
    memcpy(mu, mu2, sizeof(double) * (size_t)(N * D));
    memcpy(Sigma, S2, sizeof(double) * (size_t)(N * D * D));
    for (int64_t i = 0; i < N; ++i) logw[i] = 0.0;
    *typeless = 0;
    baseline = orc_logsumexp(N, logw);
  } else {
    for (int64_t i = 0; i < N; ++i) anc[i] = (int32_t)i;
    if (*typeless) {
      baseline = log((double)N);
      for (int64_t i = 0; i < N; ++i) logw[i] = 0.0;
      *typeless = 0;
    } else {
      baseline = orc_logsumexp(N, logw);
    }
  }
  for (int64_t i = 0; i < N; ++i) {
    const double xprev = xo[i];
    const double xnew = m->oa * xprev + m->ob + sqrt(m->oq) * zo[i];
    xo[i] = xnew;
    for (int r = 0; r < D; ++r) bt[r] = m->b[r] + m->C[r] * xprev;
    orc_kf_predict(D, m->A, bt, m->Q, mu + i * D, Sigma + i * D * D);
    const double sc = exp(m->rs * xnew);
    for (int e = 0; e < Dy * Dy; ++e) Rt[e] = sc * m->R0[e];
    double ll;
    rc = orc_kf_update(D, Dy, m->H, m->c, Rt, y, 0.0, mu + i * D, Sigma + i * D * D, &ll);
    if (rc) goto done;
    logw[i] += ll;
  }
  {
    const double lse = orc_logsumexp(N, logw);
    *ll_inc = lse - baseline;
    for (int64_t i = 0; i < N; ++i) logw[i] -= lse;
  }
done:
  free(w);
  free(xo2);
  free(mu2);
  free(S2);
  free(bt);
  free(Rt);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* Path storage (Jacob, Murray & Rubenthaler 2015)                           */
/* ------------------------------------------------------------------------- */
/*
 * Sequential restatement of ParallelParticleTree (GF/ext/CUDAExt.jl:146-262) in the
 * serial order its parallel loops stand for.  One pool of M nodes holds the roots
 * (time 0) and every later generation; each node has a state (D doubles), a parent
 * (-1 for a root) and an offspring count.  A node is free when its count is 0 and it is
 * not a current leaf.
 *
 * The reference's prune loop (update_offspring!, CUDAExt.jl:175-185) decrements without
 * atomics; its intended result, which this restates, is the serial one: a childless leaf
 * frees itself, and every ancestor whose count reaches zero is freed in turn.
 */
typedef struct {
  int64_t N, M;
  int D;
  int depth;
  double* states;    /* states[k*D + d] */
  int32_t* parents;  /* -1 at roots */
  int32_t* offspring;
  int32_t* leaves;   /* current leaf node of each particle */
} orc_ptree;

/* ParallelParticleTree(states, M), CUDAExt.jl:152-164: roots in slots 0..N-1. */
orc_ptree* orc_ptree_create(int64_t N, int D, int64_t M, const double* x0) {
  if (N < 1 || D < 1 || M < N) return NULL;
  orc_ptree* t = (orc_ptree*)calloc(1, sizeof(orc_ptree));
  t->N = N;
  t->M = M;
  t->D = D;
  t->states = (double*)calloc((size_t)M * D, sizeof(double));
  t->parents = (int32_t*)malloc(sizeof(int32_t) * M);
  t->offspring = (int32_t*)calloc((size_t)M, sizeof(int32_t));
  t->leaves = (int32_t*)malloc(sizeof(int32_t) * N);
  for (int64_t k = 0; k < M; ++k) t->parents[k] = -1;
  memcpy(t->states, x0, sizeof(double) * N * D);
  for (int64_t i = 0; i < N; ++i) t->leaves[i] = (int32_t)i;
  return t;
}

void orc_ptree_free(orc_ptree* t) {
  if (!t) return;
  free(t->states);
  free(t->parents);
  free(t->offspring);
  free(t->leaves);
  free(t);
}

/* expand!, CUDAExt.jl:213-227: the pool doubles, new slots are free. */
static void ptree_expand(orc_ptree* t) {
  const int64_t M2 = 2 * t->M;
  t->states = (double*)realloc(t->states, sizeof(double) * M2 * t->D);
  t->parents = (int32_t*)realloc(t->parents, sizeof(int32_t) * M2);
  t->offspring = (int32_t*)realloc(t->offspring, sizeof(int32_t) * M2);
  for (int64_t k = t->M; k < M2; ++k) {
    t->parents[k] = -1;
    t->offspring[k] = 0;
    for (int d = 0; d < t->D; ++d) t->states[k * t->D + d] = 0.0;
  }
  t->M = M2;
}

/*
 * insert!(tree, states, ancestors), CUDAExt.jl:187-211, ancestors 0-based:
 *   b = leaves[anc]; offspring[leaves] = counts(anc); prune; expand when fewer than N
 *   slots are free; new leaf i = the i-th free slot (searchsortedfirst over the cumsum of
 *   the free flags); parents[new] = b; states[new] = x.
 * Returns the number of pool doublings.
 */
int orc_ptree_insert(orc_ptree* t, const double* x, const int32_t* anc) {
  const int64_t N = t->N;
  int32_t* b = (int32_t*)malloc(sizeof(int32_t) * N);
  int32_t* cnt = (int32_t*)calloc((size_t)N, sizeof(int32_t));
  for (int64_t i = 0; i < N; ++i) {
    b[i] = t->leaves[anc[i]];
    cnt[anc[i]] += 1;
  }
  for (int64_t j = 0; j < N; ++j) t->offspring[t->leaves[j]] = cnt[j];
  for (int64_t j = 0; j < N; ++j) {
    int32_t k = t->leaves[j];
    while (k >= 0 && t->offspring[k] == 0) {
      k = t->parents[k];
      if (k >= 0) t->offspring[k] -= 1;
    }
  }
  int64_t nfree = 0;
  for (int64_t k = 0; k < t->M; ++k) nfree += (t->offspring[k] == 0);
  int grew = 0;
  if (nfree < N) {
    ptree_expand(t);
    ++grew;
  }
  int64_t i = 0;
  for (int64_t k = 0; k < t->M && i < N; ++k) {
    if (t->offspring[k] != 0) continue;
    t->leaves[i] = (int32_t)k;
    t->parents[k] = b[i];
    memcpy(t->states + k * t->D, x + i * t->D, sizeof(double) * t->D);
    ++i;
  }
  ++t->depth;
  free(b);
  free(cnt);
  return grew;
}

/* get_ancestry(tree, T), CUDAExt.jl:232-245: paths[(i*(depth+1) + s)*D + d]. */
void orc_ptree_ancestry(const orc_ptree* t, double* paths) {
  const int T = t->depth;
  for (int64_t i = 0; i < t->N; ++i) {
    int32_t k = t->leaves[i];
    for (int s = T; s >= 0; --s) {
      memcpy(paths + ((size_t)i * (T + 1) + s) * t->D, t->states + (size_t)k * t->D,
             sizeof(double) * t->D);
      if (s > 0) k = t->parents[k];
    }
  }
}

/* The pool (M entries) and the leaves, for comparison with the device tree. */
int64_t orc_ptree_size(const orc_ptree* t) { return t->M; }

void orc_ptree_export(const orc_ptree* t, int32_t* parents, int32_t* offspring,
                      int32_t* leaves) {
  if (parents) memcpy(parents, t->parents, sizeof(int32_t) * t->M);
  if (offspring) memcpy(offspring, t->offspring, sizeof(int32_t) * t->M);
  if (leaves) memcpy(leaves, t->leaves, sizeof(int32_t) * t->N);
}
