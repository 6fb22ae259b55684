// pf_common.h -- tiles, per-tile weight partials and the finish kernels shared by the
// particle filters and the resampler.
//
// Particles are processed in tiles of kTile = 1024 (256 threads x 4, lane-striped: item k
// of thread t is particle tile0 + 256 k + t, so every load and store of a wave is one
// contiguous 256-byte run).  At the end of a filter step every tile b leaves a partial:
//   m0_b, s0_b  max and sum exp(nl - m0) of the pre-update normalised log-weights
//               (the predict baseline, GF/src/algorithms/particles.jl:71-73);
//   m1_b        max of the updated log-weights l (in the filter's type T);
//   tau_b       sum_i e_i in FP64 with e_i = exp(l_i - m1_b) evaluated in T;
//   s2_b        sum_i e_i^2 in FP64.
// The finish kernel (one block up to 1024 tiles, two group passes beyond) folds them in a
// fixed order into
//   lse = M + log S (M = max m1_b, S = sum tau_b exp(m1_b - M)),  ll = lse - baseline
//   (marginalise!, GF/src/containers.jl:160-173),  ESS = S^2 / sum s2_b exp(2(m1_b - M))
//   (GF/src/resamplers.jl:179-183),
// and the resampling weights of the next step: w_i = e_i * c_b with c_b = exp(m1_b - M)/S
// (softmax of the log-weights, containers.jl:140-141), whose tile sums c_b tau_b it scans
// into tile prefixes.  The resampler recomputes e_i from (l_i, m1_b) with the same FP32
// operations, so the weights it scans are exactly the ones these sums describe.
#pragma once
#include "device_common.h"
#include "internal.h"

namespace ssm {

constexpr int kTileThreads = 256;
constexpr int kTileItems = 4;
constexpr int kTile = kTileThreads * kTileItems;
constexpr int kPartial = 5;  // doubles per tile partial (field-major)

// Per-filter device scratch for tiles (all sized by the tile count nb).
struct TileScratch {
  double* partials;  // nb * kPartial
  double* blk_m;     // nb: m1_b (exactly representable in T): the partials' m1 row
  double* blk_c;     // nb: c_b
  double* prefix;    // nb + 1: exclusive tile prefixes of the weights, prefix[nb] = total
  double* groups;    // kGroupRec per group of kFinishGroup tiles (finishes of > 1024 tiles)
};

constexpr int kFinishGroup = 1024;  // tiles one finish block folds
constexpr int kGroupRec = 8;        // doubles per group record

// Exponential of the weights, e = 2^(v log2 e) on the hardware exp2 (v <= 0 here, so the
// relative error stays below 6e-6 even at v = -88; at v = 0 it is exactly 1).  The step
// kernel's partials, the finish's scales and the resampler's rescan all call this one
// function, so the weights the resampler scans are exactly the ones the sums describe.
__device__ __forceinline__ float exp_w(float v) {
  return __builtin_amdgcn_exp2f(v * 1.44269504088896341f);
}
__device__ __forceinline__ double exp_w(double v) { return exp(v); }

// Writes tile b's partial from each thread's 4 items.  nl: pre-update normalised
// log-weights (ignored when `fresh`), l: updated log-weights, valid: item < N.
// Partials are stored field-major: field k of tile b at partials[k * nb + b].
// Wave reductions run on DPP (fixed association); the block needs two barriers.
template <typename T>
__device__ void tile_partial(const T nl[kTileItems], const T l[kTileItems],
                             const bool valid[kTileItems], bool fresh, double* partials,
                             int nb, int b) {
  constexpr int NW = kTileThreads / kWave;
  __shared__ T lds_m1[NW];
  __shared__ double lds_r[4][NW];
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  // tile max of the updated log-weights
  T mt = (T)-INFINITY;
#pragma unroll
  for (int k = 0; k < kTileItems; ++k)
    if (valid[k]) mt = l[k] > mt ? l[k] : mt;
  mt = wave_max_dpp(mt);
  if (lane == 0) lds_m1[w] = mt;
  // baseline partial (pre-update): per-wave (max, sum exp(nl - max)); `fresh` is uniform
  T m0 = (T)-INFINITY, s0 = (T)0;
  if (!fresh) {
#pragma unroll
    for (int k = 0; k < kTileItems; ++k)
      if (valid[k]) m0 = nl[k] > m0 ? nl[k] : m0;
    const T mw = wave_max_dpp(m0);
    T s = (T)0;
    if (mw != (T)-INFINITY) {
#pragma unroll
      for (int k = 0; k < kTileItems; ++k)
        if (valid[k]) s += exp_t(nl[k] - mw);
    }
    s0 = wave_sum_dpp(s);
    m0 = mw;
  }
  __syncthreads();
  T m1 = lds_m1[0];
#pragma unroll
  for (int k = 1; k < NW; ++k) m1 = lds_m1[k] > m1 ? lds_m1[k] : m1;
  double tau = 0.0, s2 = 0.0;
  if (m1 != (T)-INFINITY) {
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
      if (valid[k]) {
        const T e = exp_w(l[k] - m1);
        tau += (double)e;
        s2 += (double)e * (double)e;
      }
    }
  }
  tau = wave_sum_dpp(tau);
  s2 = wave_sum_dpp(s2);
  if (lane == 0) {
    lds_r[0][w] = tau;
    lds_r[1][w] = s2;
    lds_r[2][w] = (double)m0;
    lds_r[3][w] = (double)s0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tt = lds_r[0][0], ss = lds_r[1][0];
#pragma unroll
    for (int k = 1; k < NW; ++k) {
      tt += lds_r[0][k];
      ss += lds_r[1][k];
    }
    // a fresh step has no baseline (the finish uses log N): the empty partial
    LseAcc<double> b0{-INFINITY, 0.0, 0.0};
    if (!fresh) {
      b0 = LseAcc<double>{lds_r[2][0], lds_r[3][0], 0.0};
#pragma unroll
      for (int k = 1; k < NW; ++k)
        b0 = lse_combine(b0, LseAcc<double>{lds_r[2][k], lds_r[3][k], 0.0});
    }
    partials[b] = b0.m;
    partials[nb + b] = b0.s;
    partials[2 * nb + b] = (double)m1;
    partials[3 * nb + b] = tt;
    partials[4 * nb + b] = ss;
  }
}

// Partials of stored log-weights (filters whose step kernel is not tile-shaped): one block
// per tile.  nl_pre holds the pre-update normalised log-weights; null means there is no
// baseline to form (imported weights).
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
    k_tile_partials(int N, const T* __restrict__ l, const T* __restrict__ nl_pre,
                    const BfScalars* sc, double threshold_n, double* partials) {
  const bool rs = threshold_n >= sc->ess;
  const bool fresh = nl_pre == nullptr || rs || sc->typeless;
  const int tile0 = blockIdx.x * kTile;
  T nl[kTileItems], lv[kTileItems];
  bool valid[kTileItems];
#pragma unroll
  for (int k = 0; k < kTileItems; ++k) {
    const int i = tile0 + k * kTileThreads + threadIdx.x;
    valid[k] = i < N;
    lv[k] = valid[k] ? l[i] : (T)0;
    nl[k] = (valid[k] && !fresh) ? nl_pre[i] : (T)0;
  }
  tile_partial<T>(nl, lv, valid, fresh, partials, gridDim.x, blockIdx.x);
}

// A folded step as the scalars need it (fold_scalars): the maxima M1 (updated) and M0
// (baseline), the scaled sums S, S2 and S0, the total of the prefixes, and whether every
// weight is zero; per-thread tile fields for the folds that keep them in this form.
template <int CH>
struct TileFold {
  int lo, hi;
  double m1[CH], c[CH], pre[CH];
  double M1, S, S2, M0, S0, tot;
  bool dead;
};

// ESS of a folded step (resamplers.jl:179-183): N for dead weights, clamped to <= N.
template <int CH>
__device__ __forceinline__ double fold_ess(const TileFold<CH>& F, int N) {
  if (F.dead) return (double)N;
  double ess = (F.S * F.S) / F.S2;
  if (!(ess <= (double)N)) ess = (double)N;  // ESS <= N exactly; clamp rounding
  return ess;
}

// The scalars a folded step leaves behind (marginalise!, containers.jl:160-173; ESS,
// resamplers.jl:179-183): reads the pre-step block sc_in, returns the new one.  rs and
// fresh are the decisions the folded step was run with.
template <typename T, int CH>
__device__ __forceinline__ BfScalars fold_scalars(const TileFold<CH>& F, const BfScalars& sc_in,
                                                  int N, double logN, bool rs, bool fresh,
                                                  double* ll_steps, int32_t* rs_steps,
                                                  int record, double base = NAN) {
  // base: the predict baseline when the caller knows it (the auxiliary filter's resampled
  // steps), otherwise log N after a fresh start or the log-sum-exp of the kept weights
  const double baseline = !isnan(base) ? base : fresh ? logN : F.M0 + log(F.S0);
  const int t = sc_in.step;
  const int idx = t - 1 - sc_in.ybase;
  BfScalars s = sc_in;
  if (F.dead) {
    // every particle has zero weight: the increment is -Inf; weights restart uniform
    if (record && ll_steps) ll_steps[idx] = -INFINITY;
    if (record) s.ll_total += -INFINITY;
    s.lse = 0.0;
    s.ess = (double)N;
    s.typeless = 1;
  } else {
    const double lse = F.M1 + log(F.S);
    const double ess = fold_ess(F, N);
    s.lse = lse;
    s.mxn = (double)((T)F.M1 - (T)lse);
    s.s1 = F.S;
    s.ess = ess;
    s.typeless = 0;
    if (record) {
      const double ll = lse - baseline;
      s.ll_total += ll;
      if (ll_steps) ll_steps[idx] = ll;
    }
  }
  s.total_w = F.tot;
  if (record) {
    if (rs_steps) rs_steps[idx] = rs ? 1 : 0;
    s.last_resampled = rs ? 1 : 0;
    s.step = t + 1;
  }
  return s;
}

// ---- the fold of at most 1024 tiles by one 256-thread block ------------------------
// Thread t owns the consecutive tiles 4t..4t+3.  (Folding inside every block of the next
// step's resampler instead, which saves this kernel's launch, was measured 4x slower at
// 1024 tiles: the fold's ~700 VALU instructions per wave, repeated in every block, cost
// more than the launch.)
constexpr int kFoldThreads = 256;
constexpr int kFoldTiles = 4;
constexpr int kFoldMax = kFoldThreads * kFoldTiles;

// A thread's tile partials.  The updated weights' fields are requested before the step's
// scalars are waited for (every index is clamped in bounds), so the fold starts one memory
// round trip after the kernel's arguments; the baseline fields, which only a step that
// kept its particles uses, are requested once the scalars say so.  One CU issues every
// load of the fold and the lines it can have in flight bound this phase, so a fresh step
// (resampled, or the lazy uniform weights) reads 24 of the 40 bytes a tile.
struct Fold4In {
  double m1[kFoldTiles], tau[kFoldTiles], s2[kFoldTiles], m0[kFoldTiles], s0[kFoldTiles];
};
__device__ __forceinline__ void fold_load4(int nb, const double* __restrict__ P, Fold4In& L) {
  const int b0 = kFoldTiles * (int)threadIdx.x;
#pragma unroll
  for (int q = 0; q < kFoldTiles; ++q) {
    const int bc = min(b0 + q, nb - 1);
    L.m1[q] = P[2 * nb + bc];
    L.tau[q] = P[3 * nb + bc];
    L.s2[q] = P[4 * nb + bc];
  }
}
__device__ __forceinline__ void fold_load4_base(int nb, const double* __restrict__ P,
                                                bool fresh, Fold4In& L) {
  const int b0 = kFoldTiles * (int)threadIdx.x;
  if (fresh) {
#pragma unroll
    for (int q = 0; q < kFoldTiles; ++q) {
      L.m0[q] = -INFINITY;
      L.s0[q] = 0.0;
    }
  } else {
#pragma unroll
    for (int q = 0; q < kFoldTiles; ++q) {
      const int bc = min(b0 + q, nb - 1);
      L.m0[q] = P[bc];
      L.s0[q] = P[nb + bc];
    }
  }
}

struct Fold4 {
  double m1[kFoldTiles], c[kFoldTiles], pre[kFoldTiles];  // tiles 4t..4t+3
  double M1, S, S2, M0, S0, tot;
  bool dead;
};

// Folds the partials of nb <= kFoldMax tiles: maxima, the scaled sums (exp(m1_b - M1) in T)
// and the exclusive prefixes of v_b = c_b tau_b, nondecreasing in tile order.  In a wave,
// thread t's tiles take e_q = v_0 + ... + v_{q-1} in order; X_t, the exclusive tree scan of
// the thread sums made monotone by an exact running max, is the thread's base, and tile q's
// relative prefix is min(X_t + e_q, X_{t+1}) (X_64 = E, the wave's closing value max(total,
// X_63)): nondecreasing in q, and never above the next thread's first.  The waves are
// chained by off_{w+1} = off_w + E_w, so no prefix exceeds the next tile's; tot is the last
// offset, at least every prefix.  The baseline partials only count when the folded step
// did not start fresh.  Three barriers.
//
// NT = 64 (one wave: nb <= 4 x 64) combines as the four-wave block does, with the three
// missing waves' values standing in as constants (maxima -inf, sums and closing values 0),
// so both forms give the same bits: with nb <= 256 those are exactly what waves 1-3 hold.
template <typename T, int NT = kFoldThreads>
__device__ __forceinline__ void fold_step4(int nb, const Fold4In& L, bool fresh, Fold4& F) {
  constexpr int NW = NT / 64, R = kFoldTiles, NV = kFoldThreads / 64;
  static_assert(NV == 4 && (NW == NV || NW == 1), "one wave or the four-wave block");
  __shared__ double red[5][NV];
  __shared__ double wend[NV];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if constexpr (NW == 1) {
    if (lane >= 1 && lane < NV) {
      red[0][lane] = -INFINITY;
      red[1][lane] = -INFINITY;
      red[2][lane] = 0.0;
      red[3][lane] = 0.0;
      red[4][lane] = 0.0;
      wend[lane] = 0.0;
    }
  }
  double m0[R], s0[R], tau[R], s2[R];
  double M1 = -INFINITY, M0 = -INFINITY;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const bool in = R * t + q < nb;
    m0[q] = (in && !fresh) ? L.m0[q] : -INFINITY;
    s0[q] = (in && !fresh) ? L.s0[q] : 0.0;
    F.m1[q] = in ? L.m1[q] : -INFINITY;
    tau[q] = in ? L.tau[q] : 0.0;
    s2[q] = in ? L.s2[q] : 0.0;
    M1 = fmax(M1, F.m1[q]);
    M0 = fmax(M0, m0[q]);
  }
  M1 = wave_max_dpp(M1);
  M0 = wave_max_dpp(M0);
  if (lane == 0) {
    red[0][wv] = M1;
    red[1][wv] = M0;
  }
  __syncthreads();
  M1 = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
  M0 = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
  double S = 0.0, S2 = 0.0, S0 = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    F.c[q] = (F.m1[q] != -INFINITY) ? (double)exp_w((T)(F.m1[q] - M1)) : 0.0;
    S += tau[q] * F.c[q];
    S2 += s2[q] * F.c[q] * F.c[q];
    if (!fresh && m0[q] != -INFINITY) S0 += s0[q] * (double)exp_w((T)(m0[q] - M0));
  }
  S = wave_sum_dpp(S);
  S2 = wave_sum_dpp(S2);
  S0 = wave_sum_dpp(S0);
  if (lane == 0) {
    red[2][wv] = S;
    red[3][wv] = S2;
    red[4][wv] = S0;
  }
  __syncthreads();
  S = ((red[2][0] + red[2][1]) + red[2][2]) + red[2][3];
  S2 = ((red[3][0] + red[3][1]) + red[3][2]) + red[3][3];
  S0 = ((red[4][0] + red[4][1]) + red[4][2]) + red[4][3];
  const bool dead = !(M1 > -INFINITY) || !(S > 0.0);
  const double invS = dead ? 0.0 : 1.0 / S;
  double e[R], run = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    F.c[q] *= invS;
    e[q] = run;
    run += F.c[q] * tau[q];
  }
  const double incl = wave_scan_dpp(run);
  const double X = wave_scan_max_dpp(dpp_d<kDppWaveShr1>(0.0, incl));  // lane 0: 0
  const double E = fmax(lane63(incl), lane63(X));
  const double Xn = dpp_d<kDppWaveShl1>(E, X);  // the next lane's base; lane 63: E
  if (lane == 0) wend[wv] = E;
  __syncthreads();
  double off = 0.0, mine = 0.0;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if (k == wv) mine = off;
    off += wend[k];
  }
#pragma unroll
  for (int q = 0; q < R; ++q) F.pre[q] = mine + fmin(X + e[q], Xn);
  F.M1 = M1;
  F.S = S;
  F.S2 = S2;
  F.M0 = M0;
  F.S0 = S0;
  F.tot = off;
  F.dead = dead;
}

// The decision a folded step ran with and, for the auxiliary particle filter, its predict
// baseline.  aux (null for the plain filters) holds the fold of the look-ahead weights
// log w_i + log eta_i: the step resampled iff threshold N >= their ESS, and after a resample
// the baseline is log N - (LSE(log w + log eta) - LSE(log w)) (construct_new_state,
// GF/src/resamplers.jl:126-138, plus the LSE(-log eta[a]) that predict adds and the
// correction subtracts), LSE(log w) being log N for the lazy uniform weights and 0 for
// normalised ones.
struct StepDecision {
  bool rs, fresh;
  double base;
};
__device__ __forceinline__ StepDecision step_decision(const BfScalars& in, const BfScalars* aux,
                                                     double threshold_n, double logN) {
  StepDecision d;
  d.rs = threshold_n >= (aux ? aux->ess : in.ess);
  d.fresh = d.rs || in.typeless;
  d.base = (aux && d.rs) ? logN - (aux->lse - (in.typeless ? logN : 0.0)) : NAN;
  return d;
}

// The folded step as the scalars need it.
__device__ __forceinline__ TileFold<1> fold_summary(const Fold4& F) {
  TileFold<1> G;
  G.M1 = F.M1;
  G.S = F.S;
  G.S2 = F.S2;
  G.M0 = F.M0;
  G.S0 = F.S0;
  G.tot = F.tot;
  G.dead = F.dead;
  return G;
}

// The finish of nb <= 1024 tiles by one block of kFoldThreads threads: the fold, every
// tile's scale and prefix (its maximum stays in the partials' m1 row, which blk_m points
// at), the total and the new scalars (sc_in may equal sc: it is read before the write).
// The single-filter finish kernel and the server's batched finish both run this function,
// so a served filter's results equal the plain path's bit for bit.
template <typename T, int NT = kFoldThreads>
__device__ __forceinline__ void pf_finish_fold4(const double* partials, const BfScalars* sc_in,
                                                int nb, int N, BfScalars* sc,
                                                double threshold_n, double logN,
                                                const TileScratch& ts, double* ll_steps,
                                                int32_t* rs_steps, int record,
                                                const BfScalars* aux) {
  Fold4In L;
  fold_load4(nb, partials, L);
  const BfScalars in = *sc_in;
  const StepDecision dec = step_decision(in, aux, threshold_n, logN);
  const bool rs = dec.rs, fresh = dec.fresh;
  fold_load4_base(nb, partials, fresh, L);
  Fold4 F;
  fold_step4<T, NT>(nb, L, fresh, F);
  const int b0 = kFoldTiles * threadIdx.x;
  if (b0 + kFoldTiles <= nb) {  // 16-byte stores (the scratch arrays are 256-byte aligned)
#pragma unroll
    for (int q = 0; q < kFoldTiles; q += 2) {
      *reinterpret_cast<double2*>(ts.prefix + b0 + q) = make_double2(F.pre[q], F.pre[q + 1]);
      *reinterpret_cast<double2*>(ts.blk_c + b0 + q) = make_double2(F.c[q], F.c[q + 1]);
    }
  } else {
#pragma unroll
    for (int q = 0; q < kFoldTiles; ++q) {
      const int b = b0 + q;
      if (b < nb) {
        ts.prefix[b] = F.pre[q];
        ts.blk_c[b] = F.c[q];
      }
    }
  }
  if (threadIdx.x == 0) {
    ts.prefix[nb] = F.tot;
    *sc = fold_scalars<T, 1>(fold_summary(F), in, N, logN, rs, fresh, ll_steps, rs_steps,
                             record, dec.base);
  }
}

// The finish kernel of nb <= 1024 tiles.  The first loads' pointers and count lead the
// arguments (preloaded into registers).
template <typename T>
__global__ void __launch_bounds__(kFoldThreads)
    k_pf_finish_fold(const double* partials, const BfScalars* sc_in, int nb, int N,
                     BfScalars* sc, double threshold_n, double logN, TileScratch ts,
                     double* ll_steps, int32_t* rs_steps, int record, const BfScalars* aux) {
  pf_finish_fold4<T>(partials, sc_in, nb, N, sc, threshold_n, logN, ts, ll_steps, rs_steps,
                     record, aux);
}

// Exclusive prefix of one value per thread across NT threads, nondecreasing in thread order
// by construction: the in-wave exclusive prefixes x are made monotone by an exact running
// max, and the wave offsets are off_{w+1} = max(off_w + tot_w, off_w + xmax_w), so no value
// of wave w exceeds the first one of wave w + 1.  *total = off_NW, at least every prefix.
// One barrier; lds_tot and lds_xmax hold NT / 64 doubles each.
template <int NT>
__device__ __forceinline__ double block_prefix_monotone(double v, double* lds_tot,
                                                        double* lds_xmax, double* total) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double incl = wave_scan_dpp(v);
  double x = wave_shfl_up(incl, 1);
  if (lane == 0) x = 0.0;
  x = wave_scan_max_dpp(x);
  if (lane == 63) {
    lds_tot[wv] = incl;
    lds_xmax[wv] = x;
  }
  __syncthreads();
  double off = 0.0, mine = 0.0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    if (k == wv) mine = off;
    off = fmax(off + lds_tot[k], off + lds_xmax[k]);
  }
  *total = off;
  return mine + x;
}

// The finish of more than 1024 tiles, in two kernels over groups of kFinishGroup tiles.
// Group pass (one block per group): the group maxima M1_g, M0_g; per tile d_b =
// exp(m1_b - M1_g) in T; the group sums A_g = sum tau_b d_b, B_g = sum s2_b d_b^2,
// C_g = sum s0_b exp(m0_b - M0_g); and the in-group monotone prefixes r_b of d_b tau_b with
// their closing value R_g (r_b goes to prefix[b] and d_b to blk_c[b] until the second pass).
template <typename T>
__global__ void __launch_bounds__(kFinishGroup)
    k_pf_finish_groups(int nb, const BfScalars* sc_in, double threshold_n, TileScratch ts,
                       const BfScalars* aux) {
  constexpr int NT = kFinishGroup, NW = NT / 64;
  __shared__ double red[5][NW];
  const bool fresh = step_decision(*sc_in, aux, threshold_n, 0.0).fresh;
  const double* P = ts.partials;
  const int b = blockIdx.x * NT + threadIdx.x;
  const bool in = b < nb;
  const int bc = in ? b : nb - 1;
  double m0 = P[bc], s0 = P[nb + bc], m1 = P[2 * nb + bc], tau = P[3 * nb + bc],
         s2 = P[4 * nb + bc];
  m0 = (in && !fresh) ? m0 : -INFINITY;
  s0 = (in && !fresh) ? s0 : 0.0;
  m1 = in ? m1 : -INFINITY;
  tau = in ? tau : 0.0;
  s2 = in ? s2 : 0.0;
  double M1 = wave_max_dpp(m1), M0 = wave_max_dpp(m0);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = M1;
    red[1][threadIdx.x >> 6] = M0;
  }
  __syncthreads();
  M1 = red[0][0];
  M0 = red[1][0];
#pragma unroll
  for (int k = 1; k < NW; ++k) {
    M1 = fmax(M1, red[0][k]);
    M0 = fmax(M0, red[1][k]);
  }
  const double d = (m1 != -INFINITY) ? (double)exp_w((T)(m1 - M1)) : 0.0;
  const double e0 = (!fresh && m0 != -INFINITY) ? s0 * (double)exp_w((T)(m0 - M0)) : 0.0;
  double A = wave_sum_dpp(tau * d), B = wave_sum_dpp(s2 * d * d), C = wave_sum_dpp(e0);
  if ((threadIdx.x & 63) == 0) {
    red[2][threadIdx.x >> 6] = A;
    red[3][threadIdx.x >> 6] = B;
    red[4][threadIdx.x >> 6] = C;
  }
  __syncthreads();
  A = red[2][0];
  B = red[3][0];
  C = red[4][0];
#pragma unroll
  for (int k = 1; k < NW; ++k) {
    A += red[2][k];
    B += red[3][k];
    C += red[4][k];
  }
  double R;
  const double r = block_prefix_monotone<NT>(d * tau, red[0], red[1], &R);
  if (in) {
    ts.prefix[b] = r;
    ts.blk_c[b] = d;
  }
  if (threadIdx.x == 0) {
    double* g = ts.groups + (size_t)blockIdx.x * kGroupRec;
    g[0] = M1;
    g[1] = M0;
    g[2] = A;
    g[3] = B;
    g[4] = C;
    g[5] = R;
  }
}

// Second pass (one block per group): every block folds the G group records in group order
// -- M1 = max M1_g, S = sum A_g exp(M1_g - M1), S2 = sum B_g exp(2 (M1_g - M1)),
// S0 = sum C_g exp(M0_g - M0) in FP64 -- then writes its tiles' scales c_b = d_b f_g and
// prefixes off_g + f_g r_b, with f_g = exp(M1_g - M1) / S and off_{g+1} = off_g + f_g R_g
// (f_g r_b <= f_g R_g, so the prefixes stay nondecreasing across groups).  Block 0 writes the
// total and the step's scalars.
template <typename T>
__global__ void __launch_bounds__(kFinishGroup)
    k_pf_finish_apply(int N, int nb, const BfScalars* sc_in, BfScalars* sc, double threshold_n,
                      double logN, TileScratch ts, double* ll_steps, int32_t* rs_steps,
                      int record, const BfScalars* aux) {
  const int G = (nb + kFinishGroup - 1) / kFinishGroup;
  const StepDecision dec = step_decision(*sc_in, aux, threshold_n, logN);
  const bool rs = dec.rs, fresh = dec.fresh;
  const double* gr = ts.groups;
  double M1 = -INFINITY, M0 = -INFINITY;
  for (int g = 0; g < G; ++g) {
    M1 = fmax(M1, gr[g * kGroupRec + 0]);
    M0 = fmax(M0, gr[g * kGroupRec + 1]);
  }
  double S = 0.0, S2 = 0.0, S0 = 0.0;
  for (int g = 0; g < G; ++g) {
    const double m1g = gr[g * kGroupRec + 0], m0g = gr[g * kGroupRec + 1];
    const double e = (m1g != -INFINITY) ? exp(m1g - M1) : 0.0;
    S += gr[g * kGroupRec + 2] * e;
    S2 += gr[g * kGroupRec + 3] * e * e;
    if (!fresh && m0g != -INFINITY) S0 += gr[g * kGroupRec + 4] * exp(m0g - M0);
  }
  const bool dead = !(M1 > -INFINITY) || !(S > 0.0);
  const double invS = dead ? 0.0 : 1.0 / S;
  double off = 0.0, off_mine = 0.0, f_mine = 0.0;
  for (int g = 0; g < G; ++g) {
    const double m1g = gr[g * kGroupRec + 0];
    const double f = ((m1g != -INFINITY) ? exp(m1g - M1) : 0.0) * invS;
    if (g == (int)blockIdx.x) {
      off_mine = off;
      f_mine = f;
    }
    off = off + f * gr[g * kGroupRec + 5];
  }
  const int b = blockIdx.x * kFinishGroup + threadIdx.x;
  if (b < nb) {
    ts.blk_c[b] = ts.blk_c[b] * f_mine;
    ts.prefix[b] = off_mine + f_mine * ts.prefix[b];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ts.prefix[nb] = off;
    TileFold<1> F;
    F.M1 = M1;
    F.S = S;
    F.S2 = S2;
    F.M0 = M0;
    F.S0 = S0;
    F.tot = off;
    F.dead = dead;
    *sc = fold_scalars<T, 1>(F, *sc_in, N, logN, rs, fresh, ll_steps, rs_steps, record,
                             dec.base);
  }
}

// Enqueues the finish: one block up to 1024 tiles, the two group passes beyond (ts.groups
// must hold a record per group).  aux: the auxiliary filter's look-ahead fold (null for the
// plain filters).
template <typename T>
inline void launch_pf_finish(hipStream_t s, int N, int nb, BfScalars* sc, double threshold_n,
                             double logN, const TileScratch& ts, double* ll_steps,
                             int32_t* rs_steps, int record, const BfScalars* aux = nullptr) {
  if (nb <= kFoldMax) {
    hipLaunchKernelGGL((k_pf_finish_fold<T>), dim3(1), dim3(kFoldThreads), 0, s,
                       (const double*)ts.partials, (const BfScalars*)sc, nb, N, sc, threshold_n,
                       logN, ts, ll_steps, rs_steps, record, aux);
  } else {
    const int G = (nb + kFinishGroup - 1) / kFinishGroup;
    hipLaunchKernelGGL((k_pf_finish_groups<T>), dim3(G), dim3(kFinishGroup), 0, s, nb, sc,
                       threshold_n, ts, aux);
    hipLaunchKernelGGL((k_pf_finish_apply<T>), dim3(G), dim3(kFinishGroup), 0, s, N, nb, sc,
                       sc, threshold_n, logN, ts, ll_steps, rs_steps, record, aux);
  }
}

}  // namespace ssm
