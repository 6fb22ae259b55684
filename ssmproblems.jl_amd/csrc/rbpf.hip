// rbpf.hip -- Rao-Blackwellised particle filter on gfx950: a scalar outer state sampled by
// a bootstrap filter, and per particle an inner linear-Gaussian state carried as an
// analytic Kalman distribution (mean + covariance).
//
// Reference (paths relative to the reference checkout):
//   RBPF             GF/src/algorithms/rbpf.jl:8-84 (initialise_particle :16-22,
//                    predict_particle :24-60, update_particle :62-84)
//   hierarchical     GF/src/models/hierarchical.jl:4-45 (prev_outer / new_outer kwargs)
//   inner KF         GF/src/algorithms/kalman.jl:42-79, kalman_gradient.jl:12-29
//   step/resample    GF/src/algorithms/particles.jl:147-162, GF/src/resamplers.jl:156-183
//   exp(outer)-scaled observation noise: GF/examples/trend-inflation/script.jl:132-134
//
// Model: x_t = oa x_{t-1} + ob + N(0, oq);  z_t = A z_{t-1} + b + C x_{t-1} + N(0, Q);
//        y_t = H z_t + c + N(0, exp(rs x_t) R0).
//
// Layout in HBM (ping-pong pairs): outer states xo[i]; inner means mu[i][D]; inner
// covariances Sig[i][D*D] (one contiguous 4 KB block per particle at D = 32, so the
// ancestor gather is one coalesced read); log-weights; ancestors.  Each wave owns one
// particle at a time: it gathers the ancestor's (x, mu, Sigma), runs the inner predict and
// Joseph update with every product on v_mfma_f32_16x16x4f32 and S factorised in FP64,
// writes the new particle and folds its log-weight into the step reduction.
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "internal.h"
#include "kf_team.h"
#include "pf_common.h"

namespace ssm {

constexpr int kRbDX = 32;  // inner state dim (padded)
constexpr int kRbDY = 16;  // observation dim (padded)
// Twelve waves (one particle each) a block and a CU: three per SIMD, so one wave's serial
// phases (the FP64 inverse, LDS round trips) overlap the others' matrix-core work.  Each
// wave's LDS is 12.3 KB, the block's parameters 7.4 KB.
constexpr int kRbWaves = 12;
constexpr int kRbThreads = 64 * kRbWaves;

struct RbModelDev {
  float A[kRbDX * kRbDX], Q[kRbDX * kRbDX], H[kRbDY * kRbDX], R0[kRbDY * kRbDY];
  float b[kRbDX], C[kRbDX], c[kRbDY], mu0[kRbDX], Sigma0[kRbDX * kRbDX];
  float oa, ob, sq, om0, so0, rs;
  int D, Dy;
};

// LDS: parameters shared by the block (swizzled row-major, kf_team.h), then one region per
// wave.
struct RbShared {
  float Q[kRbDX * kRbDX], H[kRbDY * kRbDX], R0[kRbDY * kRbDY];
  float b[kRbDX], C[kRbDX], c[kRbDY], y[kRbDY];
};
// Per wave: B0 holds Sigma, then Sigma_p^T, then the new Sigma; B1 holds A Sigma, then P =
// H Sigma_p (its first 16 rows) and S (behind it, 16 x 16), then T2^T; PT holds P^T, then
// U^T.
struct RbWave {
  float B0[kRbDX * kRbDX], B1[kRbDX * kRbDX];
  float PT[kRbDX * kRbDY], K[kRbDX * kRbDY];
  float mu[kRbDX], muh[kRbDX], z[kRbDY];
};

struct RbStepParams {
  int N, nblocks;
  const float* __restrict__ xo_in;
  float* __restrict__ xo_out;
  const float* __restrict__ mu_in;
  float* __restrict__ mu_out;
  const float* __restrict__ Sig_in;
  float* __restrict__ Sig_out;
  float* __restrict__ logw;
  float* __restrict__ nl_pre;  // pre-update normalised log-weights (for the baseline)
  const int32_t* __restrict__ anc;
  const RbModelDev* __restrict__ model;
  const double* __restrict__ ys;
  BfScalars* sc;
  double threshold_n, logN;
  uint64_t seed;
  const float* __restrict__ noise;  // optional outer noise (parity mode)
  int32_t* __restrict__ err;  // set to 1 if an innovation covariance is not PD
};

// One wave per particle (grid-stride): gather the ancestor's (x, mu, Sigma), then the inner
// predict and Joseph update in the phases of the batched Kalman kernel (kf.hip), within the
// wave: T1 = A Sigma and Sigma_p = T1 A' + Q with A's two row blocks in registers; P' and P;
// S = P H' + R_t (exactly symmetric); S^-1, log det S and z' S^-1 z in FP64; K = P' S^-1;
// T2 = Sigma_p - P' K'; U = R_t K' - H T2; Sigma = Symmetric(:U)(T2 + K U).  R_t =
// exp(rs x_t) R0 is R0 scaled in the products.
__global__ void __launch_bounds__(kRbThreads) __attribute__((amdgpu_waves_per_eu(3)))
k_rbpf_step(RbStepParams p) {
  extern __shared__ __align__(16) unsigned char rb_lds[];
  RbShared& sh = *reinterpret_cast<RbShared*>(rb_lds);
  const int wave = threadIdx.x >> 6, lane0 = threadIdx.x & 63;
  RbWave& w = *reinterpret_cast<RbWave*>(rb_lds + sizeof(RbShared) + wave * sizeof(RbWave));
  float* const Sm = w.B1 + kRbDY * kRbDX;  // S, behind P
  const RbModelDev& m = *p.model;
  const int D0 = m.D, Dy0 = m.Dy;
  for (int e = threadIdx.x; e < kRbDX * kRbDX; e += kRbThreads)
    sh.Q[swz<kRbDX>(e / kRbDX, e % kRbDX)] = m.Q[e];
  for (int e = threadIdx.x; e < kRbDY * kRbDX; e += kRbThreads)
    sh.H[swz<kRbDX>(e / kRbDX, e % kRbDX)] = m.H[e];
  for (int e = threadIdx.x; e < kRbDY * kRbDY; e += kRbThreads)
    sh.R0[swz<kRbDY>(e / kRbDY, e % kRbDY)] = m.R0[e];
  for (int e = threadIdx.x; e < kRbDX; e += kRbThreads) {
    sh.b[e] = m.b[e];
    sh.C[e] = m.C[e];
  }
  BfScalars* sc = p.sc;
  const bool rs = p.threshold_n >= sc->ess;
  const bool fresh = rs || sc->typeless;
  const float lse = (float)sc->lse;
  const int t = sc->step;
  const double* yrow = p.ys + (size_t)(t - 1 - sc->ybase) * Dy0;
  for (int q = threadIdx.x; q < kRbDY; q += kRbThreads) {
    sh.c[q] = m.c[q];
    sh.y[q] = q < Dy0 ? (float)yrow[q] : 0.f;
  }
  // this wave's Sigma is zero outside the D x D block (every later write keeps it so)
  for (int e = lane0; e < kRbDX * kRbDX; e += 64) w.B0[e] = 0.f;
  // A's row blocks in the matrix-core operand layout: areg[r][kb][u] = A[16 r + li][16 kb + 4 lk + u]
  f4v areg[2][2];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
      areg[r][kb] = *reinterpret_cast<const f4v*>(m.A + (16 * r + (lane0 & 15)) * kRbDX +
                                                  16 * kb + 4 * (lane0 >> 4));
  __syncthreads();

  const double log2pi = 1.8378770664093454835606594728112;
  const int G = gridDim.x * kRbWaves;
  // The ancestor's (x, mu, Sigma) of the wave's next particle are loaded into registers
  // while the current one computes (lane l holds Sigma entries l + 64 k, k < 16): the
  // gather's HBM round trips leave the critical path.
  constexpr int kSigRegs = kRbDX * kRbDX / 64;
  float sig[kSigRegs], mu_pre, x_pre;
  auto prefetch = [&](int ip) {
    const int ap = rs ? p.anc[ip] : ip;
    const float* sg = p.Sig_in + (size_t)ap * D0 * D0;
#pragma unroll
    for (int k = 0; k < kSigRegs; ++k) {
      const int e = lane0 + 64 * k;
      sig[k] = e < D0 * D0 ? sg[e] : 0.f;
    }
    mu_pre = lane0 < D0 ? p.mu_in[(size_t)ap * D0 + lane0] : 0.f;
    x_pre = p.xo_in[ap];
  };
  const int i0 = blockIdx.x * kRbWaves + wave;
  if (i0 < p.N) prefetch(i0);
  for (int i = i0; i < p.N; i += G) {
    // The lane and shape indices pass through empty asm each particle, so the compiler
    // recomputes the LDS addresses instead of keeping them all live across the loop.
    int lane = lane0, D = D0, Dy = Dy0;
    asm volatile("" : "+v"(lane));
    asm volatile("" : "+s"(D), "+s"(Dy));
    const int li = lane & 15, lk = lane >> 4;
    const int dxp = (D + 15) & ~15, tn = dxp >> 4;
    // entry e = lane + 64 k of the D x D block is (r, c) = (e / D, e % D), advanced by
    // 64 = q D + m each k
    const int qd = 64 / D, md = 64 - qd * D;
    const float xprev = x_pre;
    const float nl = fresh ? 0.f : p.logw[i] - lse;
    float zo;
    if (p.noise) {
      zo = p.noise[i];
    } else {
      const U4 u = philox4x32_10((uint32_t)i, (uint32_t)t, kStreamOuter, 0u, (uint32_t)p.seed,
                                 (uint32_t)(p.seed >> 32));
      float z1;
      box_muller_f32(u.x, u.y, zo, z1);
    }
    const float xnew = m.oa * xprev + m.ob + m.sq * zo;
    const float sc_r = expf(m.rs * xnew);  // R_t = exp(rs x_t) R0
    // the ancestor's inner distribution, from the prefetched registers
    {
      int r = lane / D, c = lane - (lane / D) * D;
#pragma unroll
      for (int k = 0; k < kSigRegs; ++k) {
        if (lane + 64 * k < D * D) w.B0[swz<kRbDX>(r, c)] = sig[k];
        r += qd;
        c += md;
        if (c >= D) {
          c -= D;
          ++r;
        }
      }
    }
    if (lane < kRbDX) w.mu[lane] = mu_pre;
    wave_lds_sync();
    if (i + G < p.N) prefetch(i + G);
    // P1: T1 = A Sigma; mu_hat = A mu + b + C x_prev
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (r < tn) {
        f4v acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        f4v b00 = *reinterpret_cast<const f4v*>(w.B0 + swz<kRbDX>(li, 4 * lk));
        f4v b01 = *reinterpret_cast<const f4v*>(w.B0 + swz<kRbDX>(li, 16 + 4 * lk));
        f4v b10 = *reinterpret_cast<const f4v*>(w.B0 + swz<kRbDX>(16 + li, 4 * lk));
        f4v b11 = *reinterpret_cast<const f4v*>(w.B0 + swz<kRbDX>(16 + li, 16 + 4 * lk));
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[r][0][u], b00[u], acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[r][0][u], b10[u], acc1, 0, 0, 0);
        }
        if (tn > 1) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[r][1][u], b01[u], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[r][1][u], b11[u], acc1, 0, 0, 0);
          }
        }
        kf_store<kRbDX>(w.B1, 16 * r, 0, acc0, li, lk);
        if (tn > 1) kf_store<kRbDX>(w.B1, 16 * r, 16, acc1, li, lk);
        f4v sp = areg[r][0] * *reinterpret_cast<const f4v*>(w.mu + 4 * lk);
        if (tn > 1) sp += areg[r][1] * *reinterpret_cast<const f4v*>(w.mu + 16 + 4 * lk);
        float s = (sp[0] + sp[1]) + (sp[2] + sp[3]);
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        const int row = 16 * r + li;
        if (lk == 0) w.muh[row] = (row < D) ? sh.b[row] + sh.C[row] * xprev + s : 0.f;
      }
    }
    wave_lds_sync();
    // P2: Sigma_p = T1 A' + Q, stored transposed over Sigma; z = y - H mu_hat - c
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      if (r < tn) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          if (c < tn) {
            f4v acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
              if (kb < tn)
                acc = mfma4(*reinterpret_cast<const f4v*>(w.B1 + swz<kRbDX>(16 * r + li, 16 * kb + 4 * lk)),
                            areg[c][kb], acc);
            // Q symmetric: its tile (r, c) read through Q^T's rows
            acc += kf_load_t<kRbDX>(sh.Q, 16 * r, 16 * c, li, lk);
            kf_store_t<kRbDX>(w.B0, 16 * r, 16 * c, acc, li, lk);
          }
        }
      }
    }
    if (lane < kRbDY) {
      f4v sp = {0.f, 0.f, 0.f, 0.f};
      for (int k = 0; k < dxp; k += 4)
        sp += *reinterpret_cast<const f4v*>(sh.H + swz<kRbDX>(lane, k)) *
              *reinterpret_cast<const f4v*>(w.muh + k);
      w.z[lane] = lane < Dy ? sh.y[lane] - (sh.c[lane] + ((sp[0] + sp[1]) + (sp[2] + sp[3]))) : 0.f;
    }
    wave_lds_sync();
    // P3: P^T = Sigma_p^T H' (dxp x 16), kept as is and transposed into B1 (P = H Sigma_p)
    kf_mm_tiles<kRbDX, kRbDX>(
        w.B0, sh.H, dxp, 0, 1, tn, li, lk,
        [&](int tt, int& r0, int& c0) { r0 = 16 * tt; c0 = 0; },
        [&](int r0, int c0, f4v acc) {
          kf_store<kRbDY>(w.PT, r0, c0, acc, li, lk);
          kf_store_t<kRbDX>(w.B1, r0, c0, acc, li, lk);
        });
    wave_lds_sync();
    // P4: S = P H' + R_t, its upper triangle mirrored
    {
      f4v acc = kf_mm<kRbDX, kRbDX>(w.B1, sh.H, 0, 0, dxp, li, lk);
      acc += sc_r * kf_load_t<kRbDY>(sh.R0, 0, 0, li, lk);  // R0 symmetric
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ii = 4 * lk + e;
        if (ii <= li) {
          Sm[swz<kRbDY>(ii, li)] = acc[e];
          Sm[swz<kRbDY>(li, ii)] = acc[e];
        }
      }
    }
    wave_lds_sync();
    // P5: S^-1 in FP64, log det S, z' S^-1 z and K = P^T S^-1
    double ldet, quad;
    SpdInv inv;
    const int good = spd_inverse_wave<kRbDY>(Sm, Dy, w.z, inv, &ldet, &quad);
    if (good)
      for (int r = 0; r < tn; ++r) gain_rows<kRbDY>(inv, Dy, w.PT, w.K, r);
    wave_lds_sync();
    float lnew;
    if (good) {
      lnew = nl + (float)(-0.5 * ((double)Dy * log2pi + ldet + quad));
      // P7: T2 = Sigma_p - P^T K', stored transposed over B1; mu = mu_hat + K z
      kf_mm_tiles<kRbDY, kRbDY>(
          w.PT, w.K, kRbDY, 0, 1, tn * tn, li, lk,
          [&](int tt, int& r0, int& c0) { r0 = 16 * (tt / tn); c0 = 16 * (tt % tn); },
          [&](int r0, int c0, f4v pk) {
            kf_store_t<kRbDX>(w.B1, r0, c0, kf_load_t<kRbDX>(w.B0, r0, c0, li, lk) - pk, li, lk);
          });
      if (lane < kRbDX) {
        f4v sp = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < kRbDY; q += 4)
          sp += *reinterpret_cast<const f4v*>(w.K + swz<kRbDY>(lane, q)) *
                *reinterpret_cast<const f4v*>(w.z + q);
        w.mu[lane] = lane < D ? w.muh[lane] + ((sp[0] + sp[1]) + (sp[2] + sp[3])) : 0.f;
      }
      wave_lds_sync();
      // P8: U = R_t K' - H T2 (16 x dxp), stored transposed over P^T
      kf_mm_tiles<kRbDX, kRbDX>(
          sh.H, w.B1, dxp, 0, 1, tn, li, lk,
          [&](int tt, int& r0, int& c0) { r0 = 0; c0 = 16 * tt; },
          [&](int r0, int c0, f4v ht) {
            const f4v rk = kf_mm<kRbDY, kRbDY>(sh.R0, w.K, r0, c0, kRbDY, li, lk);
            kf_store_t<kRbDY>(w.PT, r0, c0, sc_r * rk - ht, li, lk);
          });
      wave_lds_sync();
      // P9: Sigma = Symmetric(:U)(T2 + K U): the upper tiles, mirrored
      kf_mm_tiles<kRbDY, kRbDY>(
          w.K, w.PT, kRbDY, 0, 1, tn * (tn + 1) / 2, li, lk,
          [&](int tt, int& r0, int& c0) {  // (0,0), (0,1), (1,1)
            r0 = tt == 2 ? 16 : 0;
            c0 = tt == 0 ? 0 : 16;
          },
          [&](int r0, int c0, f4v ku) {
            const f4v v = kf_load_t<kRbDX>(w.B1, r0, c0, li, lk) + ku;
            if (r0 < c0) {
              kf_store<kRbDX>(w.B0, r0, c0, v, li, lk);
              kf_store_t<kRbDX>(w.B0, r0, c0, v, li, lk);
            } else {
              const int j = c0 + li;
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                const int ii = r0 + 4 * lk + e;
                if (ii <= j) {
                  w.B0[swz<kRbDX>(ii, j)] = v[e];
                  w.B0[swz<kRbDX>(j, ii)] = v[e];
                }
              }
            }
          });
      wave_lds_sync();
    } else {
      lnew = -INFINITY;
      if (lane == 0) atomicOr(p.err, 1);
    }
    // write the particle (Sigma is exactly symmetric)
    {
      float* so = p.Sig_out + (size_t)i * D * D;
      int r = lane / D, c = lane - (lane / D) * D;
#pragma unroll
      for (int k = 0; k < kSigRegs; ++k) {
        if (lane + 64 * k < D * D) so[lane + 64 * k] = w.B0[swz<kRbDX>(r, c)];
        r += qd;
        c += md;
        if (c >= D) {
          c -= D;
          ++r;
        }
      }
    }
    if (lane < D) p.mu_out[(size_t)i * D + lane] = w.mu[lane];
    if (lane == 0) {
      p.xo_out[i] = xnew;
      p.logw[i] = lnew;
      p.nl_pre[i] = nl;
    }
    wave_lds_sync();
  }
}

// initialise_particle (rbpf.jl:16-22): x ~ N(om0, os0), inner (mu0, Sigma0).
__global__ void __launch_bounds__(256) k_rbpf_init(int N, const RbModelDev* model, float* xo,
                                                   float* mu, float* Sig, float* logw,
                                                   const float* noise, uint64_t seed) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const RbModelDev& m = *model;
  float z;
  if (noise) {
    z = noise[i];
  } else {
    const U4 u = philox4x32_10((uint32_t)i, 0u, kStreamInit, 0u, (uint32_t)seed,
                               (uint32_t)(seed >> 32));
    float z1;
    box_muller_f32(u.x, u.y, z, z1);
  }
  xo[i] = m.om0 + m.so0 * z;
  logw[i] = 0.f;
  const int D = m.D;
  for (int r = 0; r < D; ++r) mu[(size_t)i * D + r] = m.mu0[r];
  for (int e = 0; e < D * D; ++e) Sig[(size_t)i * D * D + e] = m.Sigma0[(e / D) * kRbDX + e % D];
}

}  // namespace ssm

using namespace ssm;

struct ssm_rbpf {
  ssm_ctx* ctx = nullptr;
  int N = 0, D = 0, Dy = 0;
  uint64_t seed = 0;
  uint64_t* seed_dev = nullptr;  // device copy of the key (the resampler reads it there)
  float* xo[2] = {nullptr, nullptr};
  float* mu[2] = {nullptr, nullptr};
  float* Sig[2] = {nullptr, nullptr};
  int cur = 0;
  float* logw = nullptr;
  float* nl_pre = nullptr;
  int32_t* anc = nullptr;
  RbModelDev* model = nullptr;
  BfScalars* sc = nullptr;
  int nblocks = 0;
  void* rs_scratch = nullptr;
  double* w_dbg = nullptr;
  double* ys = nullptr;
  size_t ys_cap = 0;
  double* ll_steps = nullptr;
  int32_t* rs_steps = nullptr;
  int32_t* err = nullptr;
  float* noise = nullptr;
  double* u = nullptr;
  bool initialised = false;
  double ess_uniform = 0.0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
};

namespace {

bool rb_chol(int n, std::vector<double>& A) {
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
  }
  return true;
}

int rb_upload(ssm_rbpf* rb, const ssm_rb_model* m) {
  ssm_ctx* ctx = rb->ctx;
  const int D = m->D, Dy = m->Dy;
  RbModelDev h;
  std::memset(&h, 0, sizeof h);
  // column-major inputs -> row-major padded
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      h.A[i * kRbDX + j] = (float)m->A[i + j * D];
      h.Q[i * kRbDX + j] = (float)m->Q[i + j * D];
      h.Sigma0[i * kRbDX + j] = (float)m->Sigma0[i + j * D];
    }
  for (int q = 0; q < Dy; ++q)
    for (int j = 0; j < D; ++j) h.H[q * kRbDX + j] = (float)m->H[q + j * Dy];
  for (int q = 0; q < Dy; ++q)
    for (int r = 0; r < Dy; ++r) h.R0[q * kRbDY + r] = (float)m->R0[q + r * Dy];
  for (int r = 0; r < D; ++r) {
    h.b[r] = (float)m->b[r];
    h.C[r] = (float)m->C[r];
    h.mu0[r] = (float)m->mu0[r];
  }
  for (int q = 0; q < Dy; ++q) h.c[q] = (float)m->c[q];
  if (!(m->oq >= 0.0) || !(m->os0 >= 0.0))
    return set_error(ctx, SSM_ERR_ARG, "outer variances must be >= 0");
  std::vector<double> R0(m->R0, m->R0 + Dy * Dy), Q(m->Q, m->Q + D * D);
  if (!rb_chol(Dy, R0)) return set_error(ctx, SSM_ERR_NOT_PD, "R0 is not positive definite");
  h.oa = (float)m->oa;
  h.ob = (float)m->ob;
  h.sq = (float)std::sqrt(m->oq);
  h.om0 = (float)m->om0;
  h.so0 = (float)std::sqrt(m->os0);
  h.rs = (float)m->rs;
  h.D = D;
  h.Dy = Dy;
  SSM_HIP_CHECK(ctx, hipMemcpy(rb->model, &h, sizeof h, hipMemcpyHostToDevice));
  return SSM_OK;
}

size_t rb_lds_bytes() { return sizeof(RbShared) + kRbWaves * sizeof(RbWave); }

int rb_steps_cap(ssm_rbpf* rb, size_t T) {
  if (T <= rb->ys_cap) return SSM_OK;
  (void)hipFree(rb->ys);
  (void)hipFree(rb->ll_steps);
  (void)hipFree(rb->rs_steps);
  rb->ys = nullptr;
  rb->ll_steps = nullptr;
  rb->rs_steps = nullptr;
  rb->ys_cap = 0;
  if (hipMalloc(&rb->ys, sizeof(double) * T * rb->Dy) != hipSuccess ||
      hipMalloc(&rb->ll_steps, sizeof(double) * T) != hipSuccess ||
      hipMalloc(&rb->rs_steps, sizeof(int32_t) * T) != hipSuccess)
    return set_error(rb->ctx, SSM_ERR_OOM, "rbpf: observation buffers");
  rb->ys_cap = T;
  return SSM_OK;
}

int rb_enqueue_step(ssm_rbpf* rb, hipStream_t s, double threshold, int kind,
                    const double* u_dev, const float* noise_dev, bool dbg) {
  ResampleArgs a;
  a.kind = kind;
  a.n = rb->N;
  a.weight_mode = kWeightsLogF32;
  a.w = rb->logw;
  a.u = u_dev;
  a.seed = rb->seed;
  a.seed_dev = rb->seed_dev;
  a.sc = rb->sc;
  a.threshold_n = threshold * (double)rb->N;
  a.anc = rb->anc;
  a.w_debug64 = dbg ? rb->w_dbg : nullptr;
  resample_bind_scratch(a, rb->rs_scratch, rb->N);
  int rc = launch_resample(rb->ctx, s, a);
  if (rc) return rc;
  const TileScratch ts{a.partials, a.blk_m, a.blk_c, a.prefix, a.groups};
  RbStepParams p;
  p.N = rb->N;
  p.nblocks = rb->nblocks;
  p.xo_in = rb->xo[rb->cur];
  p.xo_out = rb->xo[rb->cur ^ 1];
  p.mu_in = rb->mu[rb->cur];
  p.mu_out = rb->mu[rb->cur ^ 1];
  p.Sig_in = rb->Sig[rb->cur];
  p.Sig_out = rb->Sig[rb->cur ^ 1];
  p.logw = rb->logw;
  p.nl_pre = rb->nl_pre;
  p.anc = rb->anc;
  p.model = rb->model;
  p.ys = rb->ys;
  p.sc = rb->sc;
  p.threshold_n = a.threshold_n;
  p.logN = std::log((double)rb->N);
  p.seed = rb->seed;
  p.noise = noise_dev;
  p.err = rb->err;
  hipLaunchKernelGGL(k_rbpf_step, dim3(rb->nblocks), dim3(kRbThreads), rb_lds_bytes(), s, p);
  const int nb = (rb->N + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_tile_partials<float>, dim3(nb), dim3(kTileThreads), 0, s, rb->N,
                     (const float*)rb->logw, (const float*)rb->nl_pre, (const BfScalars*)rb->sc,
                     a.threshold_n, ts.partials);
  launch_pf_finish<float>(s, rb->N, nb, rb->sc, a.threshold_n, p.logN, ts, rb->ll_steps,
                          rb->rs_steps, 1);
  SSM_HIP_CHECK(rb->ctx, hipGetLastError());
  rb->cur ^= 1;
  return SSM_OK;
}

int rb_set_step(ssm_rbpf* rb, int t, int ybase) {
  const int32_t v[2] = {t, ybase};
  SSM_HIP_CHECK(rb->ctx, hipMemcpyAsync(&rb->sc->step, v, sizeof v, hipMemcpyHostToDevice,
                                        rb->ctx->stream));
  return SSM_OK;
}

int rb_check_err(ssm_rbpf* rb) {
  int32_t e = 0;
  SSM_HIP_CHECK(rb->ctx, hipMemcpy(&e, rb->err, sizeof e, hipMemcpyDeviceToHost));
  if (e) {
    const int32_t zero = 0;
    (void)hipMemcpy(rb->err, &zero, sizeof zero, hipMemcpyHostToDevice);
    return set_error(rb->ctx, SSM_ERR_NOT_PD,
                     "rbpf: an inner innovation covariance is not positive definite");
  }
  return SSM_OK;
}

}  // namespace

extern "C" int32_t ssm_rbpf_create(ssm_ctx* ctx, const ssm_rb_model* model, int32_t N,
                                   uint64_t seed, ssm_rbpf** out) {
  if (!ctx || !out || !model) return SSM_ERR_ARG;
  *out = nullptr;
  if (N < 1) return set_error(ctx, SSM_ERR_ARG, "ssm_rbpf_create: N must be >= 1");
  if (model->D < 1 || model->D > kRbDX || model->Dy < 1 || model->Dy > kRbDY)
    return set_error(ctx, SSM_ERR_ARG, "ssm_rbpf_create: need 1<=D<=%d, 1<=Dy<=%d", kRbDX,
                     kRbDY);
  if (!model->mu0 || !model->Sigma0 || !model->A || !model->b || !model->C || !model->Q ||
      !model->H || !model->c || !model->R0)
    return set_error(ctx, SSM_ERR_ARG, "ssm_rbpf_create: null parameter");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  ssm_rbpf* rb = new ssm_rbpf();
  rb->ctx = ctx;
  rb->N = N;
  rb->D = model->D;
  rb->Dy = model->Dy;
  rb->seed = seed;
  const int Npad = (int)(((int64_t)N + 2047) / 2048 * 2048);
  int per_cu = 0;
  const void* kstep = (const void*)k_rbpf_step;
  SSM_HIP_CHECK(ctx, hipFuncSetAttribute(kstep, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)rb_lds_bytes()));
  SSM_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kstep, kRbThreads,
                                                                  rb_lds_bytes()));
  if (per_cu < 1) per_cu = 1;
  rb->nblocks = std::max(1, std::min((N + kRbWaves - 1) / kRbWaves,
                                     per_cu * (ctx->cu_count > 0 ? ctx->cu_count : 256)));
  const size_t DD = (size_t)rb->D * rb->D;
  bool ok = true;
  for (int k = 0; k < 2 && ok; ++k)
    ok = hipMalloc(&rb->xo[k], sizeof(float) * Npad) == hipSuccess &&
         hipMalloc(&rb->mu[k], sizeof(float) * Npad * rb->D) == hipSuccess &&
         hipMalloc(&rb->Sig[k], sizeof(float) * Npad * DD) == hipSuccess;
  ok = ok && hipMalloc(&rb->logw, sizeof(float) * Npad) == hipSuccess &&
       hipMalloc(&rb->anc, sizeof(int32_t) * Npad) == hipSuccess &&
       hipMalloc(&rb->model, sizeof(RbModelDev)) == hipSuccess &&
       hipMalloc(&rb->sc, sizeof(BfScalars)) == hipSuccess &&
       hipMalloc(&rb->nl_pre, sizeof(float) * Npad) == hipSuccess &&
       hipMalloc(&rb->rs_scratch, resample_scratch_bytes(N)) == hipSuccess &&
       hipMalloc(&rb->w_dbg, sizeof(double) * Npad) == hipSuccess &&
       hipMalloc(&rb->err, sizeof(int32_t)) == hipSuccess &&
       hipMalloc(&rb->seed_dev, sizeof(uint64_t)) == hipSuccess &&
       hipEventCreate(&rb->ev0) == hipSuccess && hipEventCreate(&rb->ev1) == hipSuccess;
  if (!ok) {
    ssm_rbpf_destroy(rb);
    return set_error(ctx, SSM_ERR_OOM, "ssm_rbpf_create: allocation failed (N=%d)", N);
  }
  (void)hipMemset(rb->logw, 0, sizeof(float) * Npad);
  (void)hipMemset(rb->anc, 0, sizeof(int32_t) * Npad);
  (void)hipMemset(rb->sc, 0, sizeof(BfScalars));
  (void)hipMemset(rb->rs_scratch, 0, resample_scratch_bytes(N));
  (void)hipMemset(rb->err, 0, sizeof(int32_t));
  (void)hipMemcpy(rb->seed_dev, &rb->seed, sizeof(uint64_t), hipMemcpyHostToDevice);
  int rc = rb_upload(rb, model);
  if (!rc) rc = rb_steps_cap(rb, 1);
  if (rc) {
    ssm_rbpf_destroy(rb);
    return rc;
  }
  const double w = 1.0 / N;
  double s2 = 0.0;
  for (int i = 0; i < N; ++i) s2 += w * w;
  rb->ess_uniform = 1.0 / s2;
  *out = rb;
  return SSM_OK;
}

extern "C" int32_t ssm_rbpf_destroy(ssm_rbpf* rb) {
  if (!rb) return SSM_ERR_ARG;
  (void)hipSetDevice(rb->ctx->device);
  (void)hipStreamSynchronize(rb->ctx->stream);
  void* ptrs[] = {rb->xo[0],  rb->xo[1], rb->mu[0],      rb->mu[1],   rb->Sig[0],
                  rb->Sig[1], rb->logw,  rb->nl_pre,     rb->anc,     rb->model,
                  rb->sc,     rb->rs_scratch, rb->w_dbg, rb->ys,      rb->ll_steps,
                  rb->rs_steps, rb->err, rb->noise,      rb->u,       rb->seed_dev};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (rb->ev0) (void)hipEventDestroy(rb->ev0);
  if (rb->ev1) (void)hipEventDestroy(rb->ev1);
  delete rb;
  return SSM_OK;
}

extern "C" int32_t ssm_rbpf_initialise(ssm_rbpf* rb, const double* z) {
  if (!rb) return SSM_ERR_ARG;
  ssm_ctx* ctx = rb->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (z) {
    if (!rb->noise && hipMalloc(&rb->noise, sizeof(float) * rb->N) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "rbpf noise buffer");
    std::vector<float> h(rb->N);
    for (int i = 0; i < rb->N; ++i) h[i] = (float)z[i];
    SSM_HIP_CHECK(ctx, hipMemcpy(rb->noise, h.data(), sizeof(float) * rb->N,
                                 hipMemcpyHostToDevice));
  }
  BfScalars h;
  std::memset(&h, 0, sizeof h);
  h.s1 = (double)rb->N;
  h.ess = rb->ess_uniform;
  h.typeless = 1;
  h.step = 1;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(rb->sc, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
  rb->cur = 0;
  hipLaunchKernelGGL(k_rbpf_init, dim3((rb->N + 255) / 256), dim3(256), 0, ctx->stream, rb->N,
                     rb->model, rb->xo[0], rb->mu[0], rb->Sig[0], rb->logw,
                     z ? rb->noise : nullptr, rb->seed);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  rb->initialised = true;
  return SSM_OK;
}

extern "C" int32_t ssm_rbpf_step(ssm_rbpf* rb, int32_t t, const double* y, double threshold,
                                 int32_t resampler, const double* u, const double* z,
                                 double* ll_inc, int32_t* did_resample) {
  if (!rb) return SSM_ERR_ARG;
  ssm_ctx* ctx = rb->ctx;
  if (!rb->initialised) return set_error(ctx, SSM_ERR_STATE, "ssm_rbpf_step before initialise");
  if (!y || t < 1 || resampler < SSM_SYSTEMATIC || resampler > SSM_MULTINOMIAL)
    return set_error(ctx, SSM_ERR_ARG, "ssm_rbpf_step: bad arguments");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(rb->ys, y, sizeof(double) * rb->Dy, hipMemcpyHostToDevice, s));
  int rc = rb_set_step(rb, t, t - 1);
  if (rc) return rc;
  const double* u_dev = nullptr;
  if (u) {
    if (!rb->u && hipMalloc(&rb->u, sizeof(double) * rb->N) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "rbpf uniform buffer");
    const size_t nu =
        (resampler == SSM_SYSTEMATIC || resampler == SSM_SYSTEMATIC_MURRAY) ? 1 : rb->N;
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(rb->u, u, sizeof(double) * nu, hipMemcpyHostToDevice, s));
    u_dev = rb->u;
  }
  const float* noise_dev = nullptr;
  std::vector<float> h;
  if (z) {
    if (!rb->noise && hipMalloc(&rb->noise, sizeof(float) * rb->N) != hipSuccess)
      return set_error(ctx, SSM_ERR_OOM, "rbpf noise buffer");
    h.resize(rb->N);
    for (int i = 0; i < rb->N; ++i) h[i] = (float)z[i];
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(rb->noise, h.data(), sizeof(float) * rb->N,
                                      hipMemcpyHostToDevice, s));
    noise_dev = rb->noise;
  }
  rc = rb_enqueue_step(rb, s, threshold, resampler, u_dev, noise_dev, true);
  if (rc) return rc;
  double ll = 0.0;
  int32_t did = 0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&ll, rb->ll_steps, sizeof ll, hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&did, rb->rs_steps, sizeof did, hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (ll_inc) *ll_inc = ll;
  if (did_resample) *did_resample = did;
  return rb_check_err(rb);
}

extern "C" int32_t ssm_rbpf_run(ssm_rbpf* rb, int32_t t0, int32_t T, const double* ys,
                                double threshold, int32_t resampler, double* ll_total,
                                double* ll_steps) {
  if (!rb) return SSM_ERR_ARG;
  ssm_ctx* ctx = rb->ctx;
  if (!rb->initialised) return set_error(ctx, SSM_ERR_STATE, "ssm_rbpf_run before initialise");
  if (T < 1 || t0 < 1 || !ys || resampler < SSM_SYSTEMATIC || resampler > SSM_MULTINOMIAL)
    return set_error(ctx, SSM_ERR_ARG, "ssm_rbpf_run: bad arguments");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int rc = rb_steps_cap(rb, (size_t)T);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(rb->ys, ys, sizeof(double) * T * rb->Dy,
                                    hipMemcpyHostToDevice, s));
  const double zero = 0.0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&rb->sc->ll_total, &zero, sizeof zero,
                                    hipMemcpyHostToDevice, s));
  rc = rb_set_step(rb, t0, t0 - 1);
  if (rc) return rc;
  SSM_HIP_CHECK(ctx, hipEventRecord(rb->ev0, s));
  for (int k = 0; k < T && rc == SSM_OK; ++k)
    rc = rb_enqueue_step(rb, s, threshold, resampler, nullptr, nullptr, false);
  if (rc) return rc;
  SSM_HIP_CHECK(ctx, hipEventRecord(rb->ev1, s));
  double tot = 0.0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(&tot, &rb->sc->ll_total, sizeof tot, hipMemcpyDeviceToHost, s));
  if (ll_steps)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(ll_steps, rb->ll_steps, sizeof(double) * T,
                                      hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  float ms = 0.f;
  SSM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, rb->ev0, rb->ev1));
  rb->last_ms = ms;
  if (ll_total) *ll_total = tot;
  return rb_check_err(rb);
}

extern "C" int32_t ssm_rbpf_export(ssm_rbpf* rb, double* xo, double* mu, double* Sigma,
                                   double* logw, int32_t* ancestors, double* w) {
  if (!rb) return SSM_ERR_ARG;
  ssm_ctx* ctx = rb->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const int N = rb->N, D = rb->D;
  BfScalars h;
  SSM_HIP_CHECK(ctx, hipMemcpy(&h, rb->sc, sizeof h, hipMemcpyDeviceToHost));
  std::vector<float> b;
  if (xo) {
    b.resize(N);
    SSM_HIP_CHECK(ctx, hipMemcpy(b.data(), rb->xo[rb->cur], sizeof(float) * N, hipMemcpyDeviceToHost));
    for (int i = 0; i < N; ++i) xo[i] = b[i];
  }
  if (mu) {
    b.resize((size_t)N * D);
    SSM_HIP_CHECK(ctx, hipMemcpy(b.data(), rb->mu[rb->cur], sizeof(float) * N * D,
                                 hipMemcpyDeviceToHost));
    for (size_t e = 0; e < (size_t)N * D; ++e) mu[e] = b[e];
  }
  if (Sigma) {
    b.resize((size_t)N * D * D);
    SSM_HIP_CHECK(ctx, hipMemcpy(b.data(), rb->Sig[rb->cur], sizeof(float) * b.size(),
                                 hipMemcpyDeviceToHost));
    for (size_t e = 0; e < b.size(); ++e) Sigma[e] = b[e];
  }
  if (logw) {
    b.resize(N);
    SSM_HIP_CHECK(ctx, hipMemcpy(b.data(), rb->logw, sizeof(float) * N, hipMemcpyDeviceToHost));
    const float lse = (float)h.lse;
    for (int i = 0; i < N; ++i) logw[i] = h.typeless ? 0.0 : (double)(float)(b[i] - lse);
  }
  if (ancestors) {
    SSM_HIP_CHECK(ctx, hipMemcpy(ancestors, rb->anc, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    if (!h.last_resampled)
      for (int i = 0; i < N; ++i) ancestors[i] = i;
  }
  if (w) SSM_HIP_CHECK(ctx, hipMemcpy(w, rb->w_dbg, sizeof(double) * N, hipMemcpyDeviceToHost));
  return SSM_OK;
}

extern "C" int32_t ssm_rbpf_last_timing(ssm_rbpf* rb, double* ms) {
  if (!rb) return SSM_ERR_ARG;
  if (ms) *ms = rb->last_ms;
  return SSM_OK;
}
