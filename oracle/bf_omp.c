/*
 * bf_omp.c -- the all-cores CPU baseline of the bootstrap filter hot path (test and
 * measurement infrastructure: only bench.py's cpu_baseline leg and the tests load it; the
 * product never does).  The reference's algorithm (GF/src/algorithms/particles.jl:36-162,
 * GF/src/resamplers.jl:156-183 and :216-238, GF/src/containers.jl:160-173) in FP64 with
 * OpenMP over particles and the state in SoA planes:
 *   ESS rule on the normalised weights; systematic resampling from the FP64 cumulative
 *   weights, each thread filling the output range of its block of particles; propagation
 *   x' = A x + b + L_Q z with Philox4x32-10 + Box-Muller normals keyed by (seed, t, i);
 *   the observation log-density through L_R^{-1}; log-sum-exp, increment and normalisation.
 * Block-parallel sums change the association of the cumulative weights, so this is a
 * throughput baseline, not the parity oracle (oracle.c is).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXD 16

typedef struct {
  int D, Dy;
  double A[MAXD * MAXD], b[MAXD], LQ[MAXD * MAXD], H[MAXD * MAXD], c[MAXD],
      LRi[MAXD * MAXD], cst; /* column-major; LQ lower Cholesky of Q, LRi inverse of R's */
} bfomp_model;

/* #{k in [0, n) : u0 + k < n W} exactly (the systematic rule of resamplers.jl:229-235). */
static int64_t count_below(double W, double u0, int64_t n) {
  const double vs = (double)n * W;
  double e = ceil(vs - u0);
  if (!(e >= 0.0)) e = 0.0;
  if (e > (double)n) e = (double)n;
  int64_t m = (int64_t)e;
  while (m > 0 && !(u0 + (double)(m - 1) < vs)) --m;
  while (m < n && u0 + (double)m < vs) ++m;
  return m;
}

static inline void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
  }
}

/* Runs T steps of BF(N; threshold, Systematic) from x (SoA D x N planes, updated) with
 * normalised log-weights logw (N; NULL entries are the lazy uniform start when
 * typeless).  Returns the summed log-likelihood increments in *ll and the number of steps
 * that resampled. */
int bfomp_run(const bfomp_model* m, int64_t N, int T, const double* ys, double threshold,
              uint64_t seed, int nthreads, double* x, double* logw, int typeless,
              double* ll) {
  const int D = m->D, Dy = m->Dy;
  double* xn = (double*)malloc(sizeof(double) * (size_t)(N * D));
  double* w = (double*)malloc(sizeof(double) * (size_t)N);
  int32_t* anc = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
  double* bsum = (double*)malloc(sizeof(double) * (size_t)(nthreads + 1));
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  int n_rs = 0;
  double total = 0.0;
  omp_set_num_threads(nthreads);
  for (int t = 1; t <= T; ++t) {
    const double* y = ys + (size_t)(t - 1) * Dy;
    /* ESS of the normalised weights (containers.jl:140-141, resamplers.jl:179-183) */
    double s2 = 0.0;
    if (typeless) {
      for (int64_t i = 0; i < N; ++i) w[i] = 1.0 / (double)N;
      s2 = 1.0 / (double)N;
    } else {
#pragma omp parallel for reduction(+ : s2) schedule(static)
      for (int64_t i = 0; i < N; ++i) {
        w[i] = exp(logw[i]);
        s2 += w[i] * w[i];
      }
    }
    const int rs = threshold * (double)N >= 1.0 / s2;
    double baseline;
    if (rs) {
      ++n_rs;
      /* systematic resampling: block cumulative sums, then each block's output range */
      uint32_t c[4] = {0u, (uint32_t)t, 1u, 0u};
      philox(c, k0, k1);
      const double u0 = (double)(((((uint64_t)c[0] << 32) | c[1]) >> 11)) *
                        (1.0 / 9007199254740992.0);
#pragma omp parallel
      {
        const int th = omp_get_thread_num(), nt = omp_get_num_threads();
        const int64_t lo = N * th / nt, hi = N * (th + 1) / nt;
        double s = 0.0;
        for (int64_t i = lo; i < hi; ++i) s += w[i];
        bsum[th + 1] = s;
#pragma omp barrier
#pragma omp single
        {
          bsum[0] = 0.0;
          for (int k = 1; k <= nt; ++k) bsum[k] += bsum[k - 1];
        }
        /* this block's output slots: from the count at its first prefix to the count at
           the next block's, every slot written once (rounding leftovers go to the block's
           last particle) */
        const int64_t start = (th == 0) ? 0 : count_below(bsum[th], u0, N);
        const int64_t end = (th == nt - 1) ? N : count_below(bsum[th + 1], u0, N);
        double W = bsum[th];
        int64_t next = start;
        for (int64_t i = lo; i < hi; ++i) {
          W += w[i];
          const double vs = (double)N * W;
          while (next < end && u0 + (double)next < vs) anc[next++] = (int32_t)i;
        }
        while (next < end) anc[next++] = (int32_t)(hi > lo ? hi - 1 : lo);
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t i = 0; i < N; ++i)
          for (int d = 0; d < D; ++d) xn[(size_t)d * N + i] = x[(size_t)d * N + anc[i]];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < N; ++i) logw[i] = 0.0;
      }
      memcpy(x, xn, sizeof(double) * (size_t)(N * D));
      baseline = log((double)N);
    } else if (typeless) {
      for (int64_t i = 0; i < N; ++i) logw[i] = 0.0;
      baseline = log((double)N);
    } else {
      baseline = 0.0; /* LSE of normalised weights */
    }
    typeless = 0;
    /* predict and update */
    double mx = -INFINITY;
#pragma omp parallel for reduction(max : mx) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      double z[MAXD], xv[MAXD], xo[MAXD];
      for (int j = 0; 4 * j < D; ++j) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)t, 0u, (uint32_t)j};
        philox(c, k0, k1);
        for (int p = 0; p < 2; ++p) {
          const double u1 = (double)((c[2 * p] >> 8) + 1u) * (1.0 / 16777216.0);
          const double u2 = (double)(c[2 * p + 1] >> 8) * (1.0 / 16777216.0);
          const double r = sqrt(-2.0 * log(u1)), a = 6.283185307179586 * u2;
          if (4 * j + 2 * p < D) z[4 * j + 2 * p] = r * cos(a);
          if (4 * j + 2 * p + 1 < D) z[4 * j + 2 * p + 1] = r * sin(a);
        }
      }
      for (int d = 0; d < D; ++d) xv[d] = x[(size_t)d * N + i];
      for (int r = 0; r < D; ++r) {
        double s = m->b[r];
        for (int k = 0; k < D; ++k) s += m->A[r + k * D] * xv[k];
        for (int k = 0; k <= r; ++k) s += m->LQ[r + k * D] * z[k];
        xo[r] = s;
        x[(size_t)r * N + i] = s;
      }
      double res[MAXD], quad = 0.0;
      for (int q = 0; q < Dy; ++q) {
        double s = m->c[q];
        for (int k = 0; k < D; ++k) s += m->H[q + k * Dy] * xo[k];
        res[q] = y[q] - s;
      }
      for (int q = 0; q < Dy; ++q) {
        double v = 0.0;
        for (int k = 0; k <= q; ++k) v += m->LRi[q + k * Dy] * res[k];
        quad += v * v;
      }
      logw[i] += m->cst - 0.5 * quad;
      if (logw[i] > mx) mx = logw[i];
    }
    /* marginalise!: lse, increment, normalise */
    double se = 0.0;
#pragma omp parallel for reduction(+ : se) schedule(static)
    for (int64_t i = 0; i < N; ++i) se += exp(logw[i] - mx);
    const double lse = mx + log(se);
    total += lse - baseline;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) logw[i] -= lse;
  }
  *ll = total;
  free(xn);
  free(w);
  free(anc);
  free(bsum);
  return n_rs;
}
