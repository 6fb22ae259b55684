// pf_common.h -- the end of a particle-filter step shared by the filter kernels: reduce the
// (max, sum exp, sum exp^2) partials of the pre-update and updated log-weights across the
// block, hand them to the last block to finish, and let that block fold all partials (FP64,
// fixed order) into the log-likelihood increment, the normaliser of the stored
// log-weights and the ESS that decides the next step's resampling.
//   predict baseline  GF/src/algorithms/particles.jl:71-73 (logsumexp(logw) + ll_baseline)
//   marginalise!      GF/src/containers.jl:160-173 (ll = LSE_after - baseline)
//   ESS               GF/src/resamplers.jl:179-183 (1 / sum(w^2))
#pragma once
#include "device_common.h"
#include "internal.h"

namespace ssm {

// Call with every thread of the block.  a0: partial over the pre-update normalised
// log-weights; a1: over the updated ones (threads without particles pass lse_empty()).
// `fresh` means the pre-update weights were all zero (just resampled, or TypelessZero).
template <typename T, int NTHR>
__device__ void pf_step_finish(LseAcc<T> a0, LseAcc<T> a1, double* partials, int nblocks,
                               BfScalars* sc, bool fresh, bool rs, int t, double logN, int N,
                               double* ll_steps, int32_t* rs_steps) {
  constexpr int NW = NTHR / kWave;
  __shared__ LseAcc<double> w0[NW], w1[NW];
  __shared__ int is_last;
  a0 = wave_lse(a0);
  a1 = wave_lse(a1);
  const int wv = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (lane == 0) {
    w0[wv] = LseAcc<double>{(double)a0.m, (double)a0.s, 0.0};
    w1[wv] = LseAcc<double>{(double)a1.m, (double)a1.s, (double)a1.s2};
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    LseAcc<double> b0 = w0[0], b1 = w1[0];
#pragma unroll
    for (int k = 1; k < NW; ++k) {
      b0 = lse_combine(b0, w0[k]);
      b1 = lse_combine(b1, w1[k]);
    }
    double* P = partials + (size_t)blockIdx.x * 5;
    P[0] = b0.m;
    P[1] = b0.s;
    P[2] = b1.m;
    P[3] = b1.s;
    P[4] = b1.s2;
    __threadfence();
    is_last = (atomicAdd(&sc->cnt_step, 1u) == (uint32_t)(nblocks - 1));
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  LseAcc<double> c0 = lse_empty<double>(), c1 = lse_empty<double>();
  for (int b = threadIdx.x; b < nblocks; b += NTHR) {
    const double* P = partials + (size_t)b * 5;
    c0 = lse_combine(c0, LseAcc<double>{__builtin_nontemporal_load(P + 0),
                                        __builtin_nontemporal_load(P + 1), 0.0});
    c1 = lse_combine(c1, LseAcc<double>{__builtin_nontemporal_load(P + 2),
                                        __builtin_nontemporal_load(P + 3),
                                        __builtin_nontemporal_load(P + 4)});
  }
  c0 = wave_lse(c0);
  c1 = wave_lse(c1);
  __syncthreads();
  if (lane == 0) {
    w0[wv] = c0;
    w1[wv] = c1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    LseAcc<double> b0 = w0[0], b1 = w1[0];
#pragma unroll
    for (int k = 1; k < NW; ++k) {
      b0 = lse_combine(b0, w0[k]);
      b1 = lse_combine(b1, w1[k]);
    }
    const double baseline = fresh ? logN : b0.m + log(b0.s);
    const int idx = t - 1 - sc->ybase;
    if (b1.m == -INFINITY) {
      // Every particle has zero weight: the increment is -Inf; weights restart uniform.
      if (ll_steps) ll_steps[idx] = -INFINITY;
      sc->ll_total += -INFINITY;
      sc->lse = 0.0;
      sc->mxn = 0.0;
      sc->s1 = (double)N;
      sc->ess = (double)N;
      sc->typeless = 1;
    } else {
      const double lse_after = b1.m + log(b1.s);
      const double ll = lse_after - baseline;
      double ess = (b1.s * b1.s) / b1.s2;
      if (!(ess <= (double)N)) ess = (double)N;  // ESS <= N exactly; clamp rounding
      sc->lse = lse_after;
      sc->mxn = (double)((T)b1.m - (T)lse_after);
      sc->s1 = b1.s;
      sc->ess = ess;
      sc->ll_total += ll;
      sc->typeless = 0;
      if (ll_steps) ll_steps[idx] = ll;
    }
    if (rs_steps) rs_steps[idx] = rs ? 1 : 0;
    sc->last_resampled = rs ? 1 : 0;
    sc->step = t + 1;
    sc->cnt_step = 0u;
  }
}

}  // namespace ssm
