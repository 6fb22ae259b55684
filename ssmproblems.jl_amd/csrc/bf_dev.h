// bf_dev.h -- device code of the bootstrap-filter step, shared by the single-filter
// kernels (bf.hip) and the batched serving kernels (serve.hip).
//
// Reference (paths relative to the reference checkout):
//   predict    GF/src/algorithms/particles.jl:52-74, :117-131, :228-245
//   update     particles.jl:76-90, :133-145
//   initialise particles.jl:36-50
#pragma once
#include <cmath>
#include <type_traits>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "internal.h"
#include "pf_common.h"

namespace ssm {

constexpr int kMaxD = 16;
constexpr int kMaxDy = 16;
constexpr int kStepThreads = kTileThreads;
constexpr int kStepItems = kTileItems;
constexpr int kStepBlock = kTile;
constexpr int kPadTo = kTile;

// Offsets of the packed model constants of dimensions (D, Dy): row-major A, b, the lower
// triangle of L_Q by rows, row-major H, c, the lower triangle of L_R^-1 by rows, cst.
constexpr int kModelPacked = 64;  // entries of the packed copy
struct PackedLayout {
  int A, b, LQ, H, c, LRi, cst, total;
};
__host__ __device__ constexpr PackedLayout packed_layout(int D, int Dy) {
  PackedLayout L{};
  L.A = 0;
  L.b = D * D;
  L.LQ = L.b + D;
  L.H = L.LQ + D * (D + 1) / 2;
  L.c = L.H + Dy * D;
  L.LRi = L.c + Dy;
  L.cst = L.LRi + Dy * (Dy + 1) / 2;
  L.total = L.cst + 1;
  return L;
}

// Per-filter model constants, precomputed on the host in FP64 and stored in `T`.
// Column-major with leading dimension D (or Dy).  `packed` repeats the ones the step kernels
// hold in registers, contiguously (packed_layout; filled when they fit), so that a wave's
// scalar loads of them touch 3 cache lines at D = 4 instead of 10.
template <typename T>
struct BfModelDev {
  T packed[kModelPacked];
  T A[kMaxD * kMaxD];
  T b[kMaxD];
  T LQ[kMaxD * kMaxD];     // lower Cholesky factor of Q
  T H[kMaxDy * kMaxD];
  T c[kMaxDy];
  T LRi[kMaxDy * kMaxDy];  // inverse of the lower Cholesky factor of R
  T mu0[kMaxD];
  T L0[kMaxD * kMaxD];     // lower Cholesky factor of Sigma0
  T cst;                   // -1/2 (Dy log 2 pi + log det R)
};

// A linear-Gaussian proposal q(x' | x, y) = N(F x + g0 + K y, P) of a guided particle
// filter (GF/src/algorithms/particles.jl:175-195), with the constants of the weight
// increment log f(x' | x) - log q(x' | x, y): the inverse of Q's lower Cholesky factor and
// -1/2 (log det Q - log det P).  Column-major, leading dimension D.
template <typename T>
struct PropDev {
  T F[kMaxD * kMaxD];
  T g0[kMaxD];
  T K[kMaxD * kMaxDy];   // D x Dy
  T LP[kMaxD * kMaxD];   // lower Cholesky factor of P
  T LQi[kMaxD * kMaxD];  // inverse of the lower Cholesky factor of Q
  T dcst;
};

template <typename T>
struct StepParams {
  int N, Npad, D, Dy, nblocks;
  const T* __restrict__ xin;
  T* __restrict__ xout;
  T* __restrict__ logw;
  const int32_t* __restrict__ anc;
  const BfModelDev<T>* __restrict__ model;
  const double* __restrict__ ys;
  BfScalars* sc;
  // the Philox key in device memory (graph replays survive a re-key); required.  Next to
  // sc, so that the kernel's first argument load brings both.
  const uint64_t* seed_dev = nullptr;
  double threshold_n;
  double logN;
  uint64_t seed;  // the key as the host knows it (the step reads seed_dev)
  const T* __restrict__ noise;  // optional pre-drawn noise planes (parity mode)
  double* __restrict__ partials;
  // conditional SMC (the REF kernels): the reference trajectory, (T_ref + 1) x D row-major,
  // particle 0's ancestor when the step resamples (null: particle 0), and the ancestor array
  // particle 0's entry is written back to
  const T* __restrict__ xref;
  const int32_t* ref_anc;
  int32_t* anc_w;
  // guided filter (the GUIDED kernels): the proposal
  const PropDev<T>* __restrict__ prop = nullptr;
  // auxiliary particle filter: the fold of the look-ahead weights (null: the bootstrap
  // filter).  Its ESS decides the resample, and a resampled particle starts from the
  // log-weight -log eta of its ancestor.
  const BfScalars* aux = nullptr;
};

// Per-filter constants of the reference's transition density log N(x'; A x + b, Q) for
// ancestor sampling: the inverse of Q's lower Cholesky factor (column-major, ld D) and
// -1/2 (D log 2 pi + log det Q).
template <typename T>
struct RefModelDev {
  T LQi[kMaxD * kMaxD];
  T cst;
};

// Standard normals z[0..D) of particle i at step t from the given stream.
template <typename T, int MD>
__device__ __forceinline__ void draw_normals(uint64_t seed, uint32_t stream, uint32_t t,
                                             uint32_t i, int D, T z[MD]);
template <>
__device__ __forceinline__ void draw_normals<float, 1>(uint64_t seed, uint32_t stream,
                                                       uint32_t t, uint32_t i, int D,
                                                       float z[1]) {
  const U4 w = philox4x32_10(i, t, stream, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
  float z1;
  box_muller_f32(w.x, w.y, z[0], z1);
}
template <typename T, int MD>
__device__ __forceinline__ void draw_normals(uint64_t seed, uint32_t stream, uint32_t t,
                                             uint32_t i, int D, T z[MD]) {
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int j = 0; j < (MD + 3) / 4; ++j) {
      if (4 * j < D) {
        float q[4];
        normals4_f32(seed, stream, t, i, (uint32_t)j, q);
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (4 * j + e < MD) z[4 * j + e] = q[e];
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < (MD + 1) / 2; ++j) {
      if (2 * j < D) {
        double q[2];
        normals2_f64(seed, stream, t, i, (uint32_t)j, q);
        z[2 * j] = q[0];
        if (2 * j + 1 < MD) z[2 * j + 1] = q[1];
      }
    }
  }
}

// Read access to the model constants inside the step kernel.  REG: copied to registers
// (compile-time dimensions); otherwise read from the device struct (column-major).
template <typename T, int MD, int MDY, bool REG>
struct ModelView;

template <typename T, int MD, int MDY>
struct ModelView<T, MD, MDY, true> {
  T a[MD * MD], lq[MD * MD], bb[MD], h[MDY * MD], cc[MDY], lri[MDY * MDY];
  T cst;
  // from the packed copy (MD, MDY are the model's dimensions here)
  __device__ __forceinline__ ModelView(const BfModelDev<T>& m, int D, int Dy) {
    constexpr PackedLayout L = packed_layout(MD, MDY);
    static_assert(L.total <= kModelPacked, "packed model constants do not fit");
    const T* pk = m.packed;
#pragma unroll
    for (int r = 0; r < MD; ++r) {
      bb[r] = pk[L.b + r];
#pragma unroll
      for (int c = 0; c < MD; ++c) {
        a[r * MD + c] = pk[L.A + r * MD + c];
        lq[r * MD + c] = (c <= r) ? pk[L.LQ + r * (r + 1) / 2 + c] : (T)0;
      }
    }
#pragma unroll
    for (int q = 0; q < MDY; ++q) {
      cc[q] = pk[L.c + q];
#pragma unroll
      for (int c = 0; c < MD; ++c) h[q * MD + c] = pk[L.H + q * MD + c];
#pragma unroll
      for (int c = 0; c < MDY; ++c)
        lri[q * MDY + c] = (c <= q) ? pk[L.LRi + q * (q + 1) / 2 + c] : (T)0;
    }
    cst = pk[L.cst];
  }
  __device__ __forceinline__ T A(int r, int c) const { return a[r * MD + c]; }
  __device__ __forceinline__ T LQ(int r, int c) const { return lq[r * MD + c]; }
  __device__ __forceinline__ T b(int r) const { return bb[r]; }
  __device__ __forceinline__ T H(int q, int c) const { return h[q * MD + c]; }
  __device__ __forceinline__ T c(int q) const { return cc[q]; }
  __device__ __forceinline__ T LRi(int q, int c) const { return lri[q * MDY + c]; }
};

template <typename T, int MD, int MDY>
struct ModelView<T, MD, MDY, false> {
  const BfModelDev<T>& m;  // a copy in LDS (see k_bf_step)
  int D, Dy;
  T cst;
  __device__ __forceinline__ ModelView(const BfModelDev<T>& mm, int D_, int Dy_)
      : m(mm), D(D_), Dy(Dy_), cst(mm.cst) {}
  __device__ __forceinline__ T A(int r, int c) const { return m.A[r + c * D]; }
  __device__ __forceinline__ T LQ(int r, int c) const { return m.LQ[r + c * D]; }
  __device__ __forceinline__ T b(int r) const { return m.b[r]; }
  __device__ __forceinline__ T H(int q, int c) const { return m.H[q + c * Dy]; }
  __device__ __forceinline__ T c(int q) const { return m.c[q]; }
  __device__ __forceinline__ T LRi(int q, int c) const { return m.LRi[q + c * Dy]; }
};

// The auxiliary particle filter's look-ahead score of a state x (RepresentativeStateLookAhead
// with MeanPredictive, GF/src/algorithms/particles.jl:282-298 and :346-352): the observation
// log-density of y at the transition mean, log N(y; H (A x + b) + c, R).  The look-ahead
// kernel and the step kernel call this one function, so a resampled particle's -log eta
// is the score its ancestor was drawn with, bit for bit.
template <typename T, int MD, int MDY, class MV>
__device__ __forceinline__ T apf_log_eta(const MV& mv, int D, int Dy, const T* x, const T* y) {
  T xm[MD];
#pragma unroll
  for (int r = 0; r < MD; ++r) {
    T s = (r < D) ? mv.b(r) : (T)0;
#pragma unroll
    for (int c = 0; c < MD; ++c)
      if (r < D && c < D) s += mv.A(r, c) * x[c];
    xm[r] = s;
  }
  T res[MDY];
#pragma unroll
  for (int q = 0; q < MDY; ++q) {
    T s = (q < Dy) ? mv.c(q) : (T)0;
#pragma unroll
    for (int c = 0; c < MD; ++c)
      if (q < Dy && c < D) s += mv.H(q, c) * xm[c];
    res[q] = (q < Dy) ? y[q] - s : (T)0;
  }
  T quad = 0;
#pragma unroll
  for (int q = 0; q < MDY; ++q) {
    if (q < Dy) {
      T v = 0;
#pragma unroll
      for (int c = 0; c <= q; ++c) v += mv.LRi(q, c) * res[c];
      quad += v * v;
    }
  }
  return mv.cst - (T)0.5 * quad;
}

// Copies the model constants to LDS (runtime-dimension kernels read them many times).
// The proposal's constants with the step's mean offset g = g0 + K y: registers for
// compile-time dimensions, the device struct otherwise.
template <typename T, int MD, int MDY, bool REG>
struct PropView;

template <typename T, int MD, int MDY>
struct PropView<T, MD, MDY, true> {
  T f[MD * MD], lp[MD * MD], lqi[MD * MD], gg[MD];
  T dcst;
  __device__ __forceinline__ PropView(const PropDev<T>& m, int D, int Dy, const T* y) {
#pragma unroll
    for (int r = 0; r < MD; ++r) {
      T s = (r < D) ? m.g0[r] : (T)0;
#pragma unroll
      for (int q = 0; q < MDY; ++q)
        if (r < D && q < Dy) s += m.K[r + q * D] * y[q];
      gg[r] = s;
#pragma unroll
      for (int c = 0; c < MD; ++c) {
        const bool in = r < D && c < D;
        f[r * MD + c] = in ? m.F[r + c * D] : (T)0;
        lp[r * MD + c] = in ? m.LP[r + c * D] : (T)0;
        lqi[r * MD + c] = in ? m.LQi[r + c * D] : (T)0;
      }
    }
    dcst = m.dcst;
  }
  __device__ __forceinline__ T F(int r, int c) const { return f[r * MD + c]; }
  __device__ __forceinline__ T LP(int r, int c) const { return lp[r * MD + c]; }
  __device__ __forceinline__ T LQi(int r, int c) const { return lqi[r * MD + c]; }
  __device__ __forceinline__ T g(int r) const { return gg[r]; }
};

template <typename T, int MD, int MDY>
struct PropView<T, MD, MDY, false> {
  const PropDev<T>& m;
  int D;
  T gg[MD];
  T dcst;
  __device__ __forceinline__ PropView(const PropDev<T>& mm, int D_, int Dy, const T* y)
      : m(mm), D(D_) {
#pragma unroll
    for (int r = 0; r < MD; ++r) {
      T s = (r < D) ? m.g0[r] : (T)0;
#pragma unroll
      for (int q = 0; q < MDY; ++q)
        if (r < D && q < Dy) s += m.K[r + q * D] * y[q];
      gg[r] = s;
    }
    dcst = m.dcst;
  }
  __device__ __forceinline__ T F(int r, int c) const { return m.F[r + c * D]; }
  __device__ __forceinline__ T LP(int r, int c) const { return m.LP[r + c * D]; }
  __device__ __forceinline__ T LQi(int r, int c) const { return m.LQi[r + c * D]; }
  __device__ __forceinline__ T g(int r) const { return gg[r]; }
};

template <typename T>
__device__ __forceinline__ const BfModelDev<T>& model_to_lds(const BfModelDev<T>* g,
                                                             BfModelDev<T>* s) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(g);
  uint32_t* dst = reinterpret_cast<uint32_t*>(s);
  for (int e = threadIdx.x; e < (int)(sizeof(BfModelDev<T>) / 4); e += blockDim.x) dst[e] = src[e];
  __syncthreads();
  return *s;
}

// The fused predict + update + partial-reduction kernel.  DT/DYT > 0 fix the dimensions
// at compile time; 0 means runtime dimensions up to kMaxD / kMaxDy.  NOISE: the normals
// come from pre-drawn planes (parity mode) instead of Philox.
//
// The block-uniform `fresh` (resampled or lazy-uniform weights: the pre-update log-weights
// are all 0) selects one of two straight-line bodies, so that no load sits under a
// condition: a conditional load becomes a branch, and at every branch the compiler waits
// for all outstanding loads, which serialised the items' round trips.  A fresh step reads
// the ancestors (used when it resampled) and gathers the states; a step that kept its
// particles reads the states in place and the log-weights.
//
// REF (conditional SMC, GF/src/algorithms/particles.jl:59-66, resamplers.jl:51-65):
// particle 0 takes the reference state x_ref[t] instead of a propagated one, and when the
// step resamples its ancestor is *ref_anc (ancestor sampling) or particle 0.
//
// SINGLE: the form for one filter's latency-bound launch (inputs' wait pinned, bulk outputs
// stored at device scope); the server's batched step, whose grid covers many filters and
// whose speed occupancy sets, takes plain loads and stores.
template <typename T, int DT, int DYT, bool NOISE = false, bool REF = false, bool GUIDED = false,
          bool APF = false, bool SINGLE = true>
__device__ __forceinline__ void bf_step_tile(const StepParams<T>& p, int tile) {
  constexpr int MD = DT ? DT : kMaxD;
  constexpr int MDY = DYT ? DYT : kMaxDy;
  const int D = DT ? DT : p.D;
  const int Dy = DYT ? DYT : p.Dy;
  const int tile0 = tile * kTile;
  // Every input that depends on nothing but the arguments is requested in one batch: the
  // tile's ancestors (anc holds Npad entries, so every index is in bounds; a step that
  // keeps its particles does not use them), the filter's scalars, the key and the model
  // constants.  The empty asm statements then take the ancestors and the scalars, so the
  // kernel waits once for the batch and the compiler cannot move a load behind the
  // decision (a load behind a condition on a loaded value waits for the one before it).
  // The gather thus starts one round trip after the kernel's arguments instead of four.
  // The server's batched step (!SINGLE) reads the ancestors inside the fresh body instead.
  static_assert(SINGLE || !REF, "the batched step has no reference particle");
  int a_in[kStepItems];
  if constexpr (SINGLE) {
#pragma unroll
    for (int k = 0; k < kStepItems; ++k) a_in[k] = p.anc[tile0 + k * kStepThreads + threadIdx.x];
  }
  // Model constants: registers for compile-time dimensions (read before any store, so the
  // loads are uniform scalar loads), an LDS copy otherwise.
  const BfModelDev<T>* msrc = p.model;
  if constexpr (DT == 0) {
    __shared__ BfModelDev<T> lds_model;
    msrc = &model_to_lds<T>(p.model, &lds_model);
  }
  const ModelView<T, MD, MDY, (DT > 0)> mv(*msrc, D, Dy);
  BfScalars* sc = p.sc;
  // the auxiliary filter (APF: p.aux set) decides on its look-ahead weights' ESS
  const double ess_in = APF ? p.aux->ess : sc->ess;
  const int32_t typeless = sc->typeless;
  const double lse_in = sc->lse;
  const int t = sc->step;
  const int ybase = sc->ybase;
  const uint64_t seed = *p.seed_dev;
  // The wait is pinned for a single filter's FP32 steps without a proposal.  With FP64
  // constants, a guided filter's proposal on top of the model, or the server's kernel that
  // holds three dimension variants, holding the whole batch in registers at once costs
  // more occupancy than the early gather gains (112 to 167 VGPRs at FP64).
  if constexpr (sizeof(T) == 4 && !GUIDED && SINGLE) {
#pragma unroll
    for (int k = 0; k < kStepItems; ++k) asm volatile("" : "+v"(a_in[k]));
    asm volatile("" ::"s"(ess_in), "s"(lse_in), "s"(typeless), "s"(t), "s"(ybase), "s"(seed));
  }
  const T lse = (T)lse_in;
  const bool rs = p.threshold_n >= ess_in;
  const bool fresh = rs | (typeless != 0);  // pre-update log-weights 0 (or -log eta: apf)
  const double* yrow = p.ys + (size_t)(t - 1 - ybase) * Dy;
  // The observation: read before the body for the guided filter (its proposal depends on
  // it), otherwise once the gathers are in flight.
  T y[MDY];
  auto load_y = [&]() {
#pragma unroll
    for (int q = 0; q < MDY; ++q) y[q] = (q < Dy) ? (T)yrow[q] : (T)0;
  };
  if constexpr (GUIDED) load_y();
  // the guided filter's proposal (uniform loads, before any store)
  struct NoProp {};
  using PV = std::conditional_t<GUIDED, PropView<T, MD, MDY, (DT > 0)>, NoProp>;
  const PV pv = [&]() {
    if constexpr (GUIDED) return PV(*p.prop, D, Dy, y);
    else return PV{};
  }();
  // the reference state at t and particle 0's ancestor (uniform loads)
  T xr[MD];
  int ref_a = 0;
  if constexpr (REF) {
#pragma unroll
    for (int d = 0; d < MD; ++d) xr[d] = (d < D) ? p.xref[(size_t)t * D + d] : (T)0;
    ref_a = p.ref_anc ? *p.ref_anc : 0;
    if (tile0 + threadIdx.x == 0) {  // item 0 of thread 0 of tile 0: particle 0
      a_in[0] = ref_a;
      if (rs) p.anc_w[0] = ref_a;
    }
  }
  // each item's source particle: its ancestor when the step resampled, itself otherwise
  int src[kStepItems];
  auto set_src = [&]() {
#pragma unroll
    for (int k = 0; k < kStepItems; ++k) {
      const int i = tile0 + k * kStepThreads + threadIdx.x;
      src[k] = (rs && i < p.N) ? a_in[k] : i;
    }
  };
  if constexpr (SINGLE) set_src();

  // Entry e of state plane `plane` (Npad entries a plane).  The server's batched step
  // (!SINGLE) indexes with 32-bit offsets: its filters hold at most kMaxD x 2^18 entries,
  // and a 32-bit offset with the uniform base takes one VGPR instead of a 64-bit address's
  // two (the batched kernel: 166 to 118 VGPRs, three to four waves a SIMD).
  auto plane_at = [&](int plane, int e) -> size_t {
    if constexpr (SINGLE)
      return (size_t)plane * p.Npad + e;
    else
      return (size_t)(uint32_t)(plane * p.Npad + e);
  };
  // the normals of item k: pre-drawn planes, or Philox keyed by (seed, t, i)
  auto normals = [&](int k, T* z) {
    const int i = tile0 + k * kStepThreads + threadIdx.x;
    if constexpr (NOISE) {
#pragma unroll
      for (int d = 0; d < MD; ++d)
        if (d < D) z[d] = p.noise[(size_t)d * p.Npad + i];
    } else {
      draw_normals<T, MD>(seed, kStreamPropagate, (uint32_t)t, (uint32_t)i, D, z);
    }
  };
  auto compute_item = [&](int k, const T* xk, const T* z, T nlk) -> T {
    const int i = tile0 + k * kStepThreads + threadIdx.x;
    T xn[MD];
    [[maybe_unused]] T lpq = 0;
    if constexpr (GUIDED) {
      // draw from the proposal: x' = F x + g + L_P z; then
      // log f(x' | x) - log q(x' | x, y) = dcst - |L_Q^-1 (x' - A x - b)|^2 / 2 + |z|^2 / 2,
      // the proposal's quadratic form being |L_P^-1 (x' - F x - g)|^2 = |z|^2
      T zz = 0;
#pragma unroll
      for (int r = 0; r < MD; ++r) {
        if (r < D) {
          T s = pv.g(r);
#pragma unroll
          for (int c = 0; c < MD; ++c)
            if (c < D) s += pv.F(r, c) * xk[c];
#pragma unroll
          for (int c = 0; c <= r; ++c) s += pv.LP(r, c) * z[c];
          xn[r] = s;
          store_out<SINGLE>(&p.xout[plane_at(r, i)], s);
          zz += z[r] * z[r];
        }
      }
      T res[MD];
#pragma unroll
      for (int r = 0; r < MD; ++r) {
        if (r < D) {
          T s = mv.b(r);
#pragma unroll
          for (int c = 0; c < MD; ++c)
            if (c < D) s += mv.A(r, c) * xk[c];
          res[r] = xn[r] - s;
        }
      }
      T qf = 0;
#pragma unroll
      for (int r = 0; r < MD; ++r) {
        if (r < D) {
          T v = 0;
#pragma unroll
          for (int c = 0; c <= r; ++c) v += pv.LQi(r, c) * res[c];
          qf += v * v;
        }
      }
      lpq = pv.dcst - (T)0.5 * qf + (T)0.5 * zz;
    } else {
      // predict: x' = A x + b + L_Q z
#pragma unroll
      for (int r = 0; r < MD; ++r) {
        if (r < D) {
          T s = mv.b(r);
#pragma unroll
          for (int c = 0; c < MD; ++c)
            if (c < D) s += mv.A(r, c) * xk[c];
#pragma unroll
          for (int c = 0; c <= r; ++c) s += mv.LQ(r, c) * z[c];
          if constexpr (REF) s = (i == 0) ? xr[r] : s;  // the pinned reference particle
          xn[r] = s;
          store_out<SINGLE>(&p.xout[plane_at(r, i)], s);
        }
      }
    }
    // update: log N(y; H x' + c, R) = cst - |L_R^-1 (y - H x' - c)|^2 / 2
    T res[MDY];
#pragma unroll
    for (int q = 0; q < MDY; ++q) {
      if (q < Dy) {
        T s = mv.c(q);
#pragma unroll
        for (int c = 0; c < MD; ++c)
          if (c < D) s += mv.H(q, c) * xn[c];
        res[q] = y[q] - s;
      }
    }
    T quad = 0;
#pragma unroll
    for (int q = 0; q < MDY; ++q) {
      if (q < Dy) {
        T v = 0;
#pragma unroll
        for (int c = 0; c <= q; ++c) v += mv.LRi(q, c) * res[c];
        quad += v * v;
      }
    }
    T l;
    if constexpr (GUIDED)
      l = nlk + lpq + (mv.cst - (T)0.5 * quad);
    else
      l = nlk + (mv.cst - (T)0.5 * quad);
    store_out<SINGLE>(&p.logw[i], l);
    return l;
  };

  auto body = [&](auto fresh_c) {
    constexpr bool FRESH = decltype(fresh_c)::value;
    T nl[kStepItems], lnew[kStepItems];
    bool valid[kStepItems];
    // the kept particles' log-weights (anc and logw hold Npad entries: in bounds)
    T lw[kStepItems];
#pragma unroll
    for (int k = 0; k < kStepItems; ++k) {
      const int i = tile0 + k * kStepThreads + threadIdx.x;
      valid[k] = i < p.N;
      if constexpr (!SINGLE && FRESH) a_in[k] = p.anc[i];
      if constexpr (!FRESH) lw[k] = p.logw[i];
    }
    // an item's source particle: its ancestor when the step resampled, itself otherwise
    auto src_of = [&](int k) {
      if constexpr (SINGLE) return src[k];
      const int i = tile0 + k * kStepThreads + threadIdx.x;
      if constexpr (FRESH)
        return (rs && valid[k]) ? a_in[k] : i;
      else
        return i;
    };
#pragma unroll
    for (int k = 0; k < kStepItems; ++k) {
      if constexpr (FRESH)
        nl[k] = (T)0;
      else
        nl[k] = lw[k] - lse;
    }
    // a particle the auxiliary filter resampled starts from -log eta of its ancestor
    // (resample_ancestor, GF/src/resamplers.jl:140-143); `apf && rs` is block-uniform
    [[maybe_unused]] const bool apf_rs = FRESH && APF && rs;
    if constexpr (DT > 0) {
      // every gather, the normals while the gathers are in flight, then the arithmetic
      T x[kStepItems][MD];
#pragma unroll
      for (int k = 0; k < kStepItems; ++k) {
        const int src = src_of(k);
#pragma unroll
        for (int d = 0; d < MD; ++d) x[k][d] = p.xin[(size_t)d * p.Npad + src];
      }
      if constexpr (!GUIDED) load_y();
      T z[kStepItems][MD];
#pragma unroll
      for (int k = 0; k < kStepItems; ++k) normals(k, z[k]);
      if constexpr (FRESH && APF) {
        if (apf_rs) {
#pragma unroll
          for (int k = 0; k < kStepItems; ++k)
            nl[k] = -apf_log_eta<T, MD, MDY>(mv, D, Dy, x[k], y);
        }
      }
#pragma unroll
      for (int k = 0; k < kStepItems; ++k) lnew[k] = compute_item(k, x[k], z[k], nl[k]);
    } else {
      if constexpr (!GUIDED) load_y();
      for (int k = 0; k < kStepItems; ++k) {
        const int src = src_of(k);
        T xk[MD], z[MD];
#pragma unroll
        for (int d = 0; d < MD; ++d)
          if (d < D) xk[d] = p.xin[plane_at(d, src)];
        normals(k, z);
        if constexpr (FRESH && APF) {
          if (apf_rs) nl[k] = -apf_log_eta<T, MD, MDY>(mv, D, Dy, xk, y);
        }
        lnew[k] = compute_item(k, xk, z, nl[k]);
      }
    }
    tile_partial<T>(nl, lnew, valid, FRESH, p.partials, p.nblocks, tile);
  };
  if (fresh)
    body(std::true_type{});
  else
    body(std::false_type{});
}

// log N(x*; A x + b, Q) = cst - |L_Q^{-1} (x* - A x - b)|^2 / 2 for one particle's state x
// and a target state x* (both in registers): the transition density of ancestor sampling
// and backward simulation.
template <typename T, int MD>
__device__ __forceinline__ T transition_logpdf(int D, const BfModelDev<T>& m,
                                               const RefModelDev<T>& rm, const T x[MD],
                                               const T xs[MD]) {
  T r[MD];
#pragma unroll
  for (int a = 0; a < MD; ++a) {
    if (a < D) {
      T s = m.b[a];
#pragma unroll
      for (int c = 0; c < MD; ++c)
        if (c < D) s += m.A[a + c * D] * x[c];
      r[a] = xs[a] - s;
    }
  }
  T quad = 0;
#pragma unroll
  for (int a = 0; a < MD; ++a) {
    if (a < D) {
      T v = 0;
#pragma unroll
      for (int c = 0; c <= a; ++c) v += rm.LQi[a + c * D] * r[c];
      quad += v * v;
    }
  }
  return rm.cst - (T)0.5 * quad;
}

// initialise: x_i = mu0 + chol(Sigma0).L z_i, log-weights 0 (particles.jl:36-50).
template <typename T, int DT>
__device__ __forceinline__ void bf_init_item(int N, int Npad, int D, const BfModelDev<T>* model,
                                             T* x, T* logw, const T* noise, uint64_t seed,
                                             int i) {
  constexpr int MD = DT ? DT : kMaxD;
  if (DT) D = DT;
  if (i >= Npad) return;
  const BfModelDev<T>& m = *model;
  T z[MD];
  if (noise) {
#pragma unroll
    for (int d = 0; d < MD; ++d)
      if (d < D) z[d] = noise[(size_t)d * Npad + i];
  } else {
    draw_normals<T, MD>(seed, kStreamInit, 0u, (uint32_t)i, D, z);
  }
#pragma unroll
  for (int r = 0; r < MD; ++r) {
    if (r < D) {
      T s = m.mu0[r];
#pragma unroll
      for (int c = 0; c <= r; ++c) s += m.L0[r + c * D] * z[c];
      x[(size_t)r * Npad + i] = (i < N) ? s : (T)0;
    }
  }
  logw[i] = (T)0;
}

// ---- host-side model preparation (shared by the filter and the server) -------------
// Lower Cholesky in place (column-major), FP64.
inline bool chol_lower(int n, std::vector<double>& A) {
  for (int j = 0; j < n; ++j) {
    double d = A[j + j * n];
    for (int k = 0; k < j; ++k) d -= A[j + k * n] * A[j + k * n];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    A[j + j * n] = d;
    for (int i = j + 1; i < n; ++i) {
      double s = A[i + j * n];
      for (int k = 0; k < j; ++k) s -= A[i + k * n] * A[j + k * n];
      A[i + j * n] = s / d;
    }
    for (int i = 0; i < j; ++i) A[i + j * n] = 0.0;
  }
  return true;
}

// Inverse of a lower-triangular matrix (column-major).
inline std::vector<double> tri_inverse(int n, const std::vector<double>& L) {
  std::vector<double> X(n * n, 0.0);
  for (int j = 0; j < n; ++j) {
    X[j + j * n] = 1.0 / L[j + j * n];
    for (int i = j + 1; i < n; ++i) {
      double s = 0.0;
      for (int k = j; k < i; ++k) s += L[i + k * n] * X[k + j * n];
      X[i + j * n] = -s / L[i + i * n];
    }
  }
  return X;
}

inline int bf_check_model(ssm_ctx* ctx, const ssm_lg_model* m) {
  if (!m || m->Dx < 1 || m->Dx > kMaxD || m->Dy < 1 || m->Dy > kMaxDy)
    return set_error(ctx, SSM_ERR_ARG, "model dims must satisfy 1<=Dx<=%d, 1<=Dy<=%d", kMaxD,
                     kMaxDy);
  if (!m->mu0 || !m->Sigma0 || !m->A || !m->b || !m->Q || !m->H || !m->c || !m->R)
    return set_error(ctx, SSM_ERR_ARG, "model has a null parameter pointer");
  return SSM_OK;
}

// The step kernels' model constants from calc_params output: Cholesky factors of Q and
// Sigma0, the inverse Cholesky factor of R and the log-density constant, in FP64 then
// rounded to T.
template <typename T>
int bf_build_model(ssm_ctx* ctx, const ssm_lg_model* mdl, BfModelDev<T>* out) {
  const int D = mdl->Dx, Dy = mdl->Dy;
  BfModelDev<T>& h = *out;
  std::memset(&h, 0, sizeof h);
  std::vector<double> LQ(mdl->Q, mdl->Q + D * D), LR(mdl->R, mdl->R + Dy * Dy),
      L0(mdl->Sigma0, mdl->Sigma0 + D * D);
  if (!chol_lower(D, LQ)) return set_error(ctx, SSM_ERR_NOT_PD, "Q is not positive definite");
  if (!chol_lower(Dy, LR)) return set_error(ctx, SSM_ERR_NOT_PD, "R is not positive definite");
  if (!chol_lower(D, L0))
    return set_error(ctx, SSM_ERR_NOT_PD, "Sigma0 is not positive definite");
  const std::vector<double> LRi = tri_inverse(Dy, LR);
  double logdet = 0.0;
  for (int q = 0; q < Dy; ++q) logdet += 2.0 * std::log(LR[q + q * Dy]);
  for (int e = 0; e < D * D; ++e) {
    h.A[e] = (T)mdl->A[e];
    h.LQ[e] = (T)LQ[e];
    h.L0[e] = (T)L0[e];
  }
  for (int e = 0; e < D; ++e) {
    h.b[e] = (T)mdl->b[e];
    h.mu0[e] = (T)mdl->mu0[e];
  }
  for (int e = 0; e < Dy * D; ++e) h.H[e] = (T)mdl->H[e];
  for (int e = 0; e < Dy; ++e) h.c[e] = (T)mdl->c[e];
  for (int e = 0; e < Dy * Dy; ++e) h.LRi[e] = (T)LRi[e];
  h.cst = (T)(-0.5 * ((double)Dy * 1.8378770664093454835606594728112 + logdet));
  const PackedLayout L = packed_layout(D, Dy);
  if (L.total <= kModelPacked) {  // the same values, packed (ModelView's register copy)
    for (int r = 0; r < D; ++r) {
      h.packed[L.b + r] = h.b[r];
      for (int c = 0; c < D; ++c) h.packed[L.A + r * D + c] = h.A[r + c * D];
      for (int c = 0; c <= r; ++c) h.packed[L.LQ + r * (r + 1) / 2 + c] = h.LQ[r + c * D];
    }
    for (int q = 0; q < Dy; ++q) {
      h.packed[L.c + q] = h.c[q];
      for (int c = 0; c < D; ++c) h.packed[L.H + q * D + c] = h.H[q + c * Dy];
      for (int c = 0; c <= q; ++c) h.packed[L.LRi + q * (q + 1) / 2 + c] = h.LRi[q + c * Dy];
    }
    h.packed[L.cst] = h.cst;
  }
  return SSM_OK;
}

// ESS of the uniform weights 1/N as the reference computes it: 1 / sum(w^2), summed left
// to right in FP64.
inline double ess_uniform(int N) {
  const double w = 1.0 / (double)N;
  double s = 0.0;
  for (int i = 0; i < N; ++i) s += w * w;
  return 1.0 / s;
}

}  // namespace ssm
