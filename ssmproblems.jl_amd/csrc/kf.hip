// kf.hip -- many independent Kalman filters of uneven size in persistent gfx950 kernels.
//
// Reference (paths relative to the reference checkout):
//   filter      GF/src/GeneralisedFilters.jl:58-84 (initialise, then predict/update per t)
//   initialise  GF/src/algorithms/kalman.jl:23-26
//   predict     kalman.jl:42-49          mu = A mu + b,  Sigma = A Sigma A' + Q
//   update      kalman_gradient.jl:12-29 with kalman.jl:51-66:
//               z = y - H mu - c;  S = H Sigma H' + R;  K = Sigma H' S^-1;
//               Sigma = (I - K H) Sigma (I - K H)' + K R K'  (Joseph form), Symmetric(:U);
//               mu = mu + K z;  ll = logpdf(MvNormal(z, S), 0)
//
// Design.  Filters are sorted by cost (Dx^3) and split into three size classes; each class
// runs one persistent kernel whose workgroups pull filters from an atomic work queue and
// keep one filter's state (Sigma, H, R and the step's temporaries) in LDS and its A and Q
// in registers for all T steps, so HBM carries only the parameters once and the
// observations.  Dimensions are padded to multiples of 16 and every dense product runs on
// the matrix cores as v_mfma_f32_16x16x4f32 tiles spread over the workgroup's waves (FP32
// in, FP32 accumulate).  The innovation covariance S is inverted in FP64 by one wave
// (Gauss-Jordan; its pivots give the log-determinant), and K = P' S^-1 is a matrix-core
// product.  The Joseph form is evaluated as T2 + K U with T2 = Sigma_p (I - K H)' and
// U = R K' - H T2, the same expression associated so that no Dx x Dx product with I - K H
// is formed; only its upper tiles are computed, as Symmetric(:U) keeps the upper triangle.
//   class    Dx      Dy     waves/WG  LDS/WG
//   small    <= 16   <= 16  1         ~9 KB
//   mid      <= 32   <= 16  2         ~19 KB
//   large    <= 64   <= 32  4         ~74 KB (two workgroups a CU)
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "internal.h"
#include "kf_common.h"
#include "kf_team.h"

namespace ssm {

template <int DXP, int DYP>
struct KfLds {
  static constexpr int kB0 = 0;                // Sigma, then Sigma_p^T, then the new Sigma
  static constexpr int kB1 = kB0 + DXP * DXP;  // A Sigma, then T2^T
  static constexpr int kH = kB1 + DXP * DXP;   // H (DYP x DXP)
  static constexpr int kP = kH + DYP * DXP;    // P = H Sigma_p (DYP x DXP)
  static constexpr int kPT = kP + DYP * DXP;   // P^T, then U^T (DXP x DYP)
  static constexpr int kK = kPT + DXP * DYP;   // K (DXP x DYP)
  static constexpr int kR = kK + DXP * DYP;    // R (DYP x DYP)
  static constexpr int kS = kR + DYP * DYP;    // S, then S^-1
  static constexpr int kMu = kS + DYP * DYP;   // vectors: mu, mu_hat, b (DXP); c, z, y (DYP)
  static constexpr int kMuh = kMu + DXP, kBv = kMuh + DXP, kCv = kBv + DXP, kZ = kCv + DYP,
                       kY = kZ + DYP;
  static constexpr int kScal = kY + DYP;       // 16 words: queue index, status
  static constexpr int kFloats = kScal + 16;
  // the FP64 inverse handed from wave 0 to the others (P5 to P6) lives in B1
  static_assert(DXP == 16 || DXP * DXP >= 2 * 64 * 4 * (DYP == 32 ? 4 : 1),
                "B1 must hold the inverse's blocks");
  static constexpr size_t kBytes = (size_t)kFloats * 4;
};

// TV: the kf_loglikelihood layout (separate arrays, one parameter block per step or one
// for all steps); otherwise each filter's packed homogeneous parameters.
// Register budget: the large class (74 KB of LDS, two workgroups a CU) at 2 waves a SIMD
// (256 VGPRs), the mid class at 3 and the small one at 4, which keeps spills to a few
// words.
template <int DXP, int DYP, int NW, bool TV>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(DXP == 64 ? 2 : DXP == 32 ? 3 : 4)))
k_kf_batch(KfBatchParams p) {
  using L = KfLds<DXP, DYP>;
  constexpr int TX = DXP / 16;  // tiles per padded state dimension, at most
  static_assert(NW >= TX, "a row block of A per wave");
  static_assert(DYP <= 32, "S's upper tiles are enumerated for at most 2 x 2");
  constexpr int NT = 64 * NW;
  extern __shared__ __align__(16) float lds[];
  float* const B0 = lds + L::kB0;
  float* const B1 = lds + L::kB1;
  float* const Hm = lds + L::kH;
  float* const Pm = lds + L::kP;
  float* const PT = lds + L::kPT;
  float* const Km = lds + L::kK;
  float* const Rm = lds + L::kR;
  float* const Sm = lds + L::kS;
  float* const mu = lds + L::kMu;
  float* const muh = lds + L::kMuh;
  float* const bv = lds + L::kBv;
  float* const cv = lds + L::kCv;
  float* const zv = lds + L::kZ;
  float* const yv = lds + L::kY;
  int* const iscal = reinterpret_cast<int*>(lds + L::kScal);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li0 = lane & 15, lk0 = lane >> 4;
  while (true) {
    const int li = li0, lk = lk0;
    if (tid == 0) iscal[0] = (int)atomicAdd(p.counter, 1u);
    __syncthreads();
    const int qi = iscal[0];
    if (qi >= p.count) break;  // every wave of the workgroup reads the same index
    const int f = p.order[qi];
    const int dx0 = p.Dx[f], dy0 = p.Dy[f];
    const int dxp0 = (dx0 + 15) & ~15, dyp0 = (dy0 + 15) & ~15;
    const int tn0 = dxp0 >> 4, tny0 = dyp0 >> 4;
    const int dx = dx0, dy = dy0, dxp = dxp0, dyp = dyp0, tn = tn0;
    // This wave's row block rb of A: it computes A Sigma's tiles (rb, c) and Sigma_p's
    // tiles (r, rb) for c, r = jw, jw + nj, ... (nj waves share the row block).
    const int rb = wave % tn, jw = wave / tn, nj = (NW - rb + tn - 1) / tn;
    const int64_t nxx = (int64_t)dx * dx, nyx = (int64_t)dy * dx, nyy = (int64_t)dy * dy;
    // block b of the TV arrays, or the packed block
    auto block = [&](int64_t b, const double** gA, const double** gb, const double** gQ,
                     const double** gH, const double** gc, const double** gR) {
      if constexpr (TV) {
        *gA = p.tv_A + b * nxx;
        *gb = p.tv_b + b * dx;
        *gQ = p.tv_Q + b * nxx;
        *gH = p.tv_H + b * nyx;
        *gc = p.tv_c + b * dy;
        *gR = p.tv_R + b * nyy;
      } else {
        (void)b;
        const double* g = p.params + p.param_off[f] + dx + nxx;
        *gA = g;
        *gb = g + nxx;
        *gQ = g + nxx + dx;
        *gH = g + 2 * nxx + dx;
        *gc = g + 2 * nxx + dx + nyx;
        *gR = g + 2 * nxx + dx + nyx + dy;
      }
    };
    // areg[kb][u] = A[16 rb + li][16 kb + 4 lk + u] (the operand layout of kf_mm, which
    // serves A Sigma as the left operand and T1 A' as the right one); Q is read from the
    // parameter block (L2) where Sigma_p's tiles add it.
    f4v areg[TX];
    const double* gQ_cur = nullptr;
    // Loads one parameter block (column-major FP64) into registers and zero-padded LDS.
    auto load_params = [&](int64_t b) {
      const double *gA, *gb, *gQ, *gH, *gc, *gR;
      block(b, &gA, &gb, &gQ, &gH, &gc, &gR);
      gQ_cur = gQ;
      const int ar = 16 * rb + li;
#pragma unroll
      for (int kb = 0; kb < TX; ++kb)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int ac = 16 * kb + 4 * lk + u;
          areg[kb][u] = (ar < dx && ac < dx) ? (float)gA[ar + (int64_t)ac * dx] : 0.f;
        }
      // (consecutive threads read consecutive rows of the column-major inputs)
      for (int e = tid; e < dyp * dxp; e += NT) {
        const int i = e % dyp, j = e / dyp;
        Hm[swz<DXP>(i, j)] = (i < dy && j < dx) ? (float)gH[i + (int64_t)j * dy] : 0.f;
      }
      for (int e = tid; e < dyp * dyp; e += NT) {
        const int i = e % dyp, j = e / dyp;
        Rm[swz<DYP>(i, j)] = (i < dy && j < dy) ? (float)gR[i + (int64_t)j * dy] : 0.f;
      }
      for (int e = tid; e < dxp; e += NT) bv[e] = e < dx ? (float)gb[e] : 0.f;
      for (int e = tid; e < dyp; e += NT) cv[e] = e < dy ? (float)gc[e] : 0.f;
    };
    const double *g_mu0, *g_S0;
    if constexpr (TV) {
      g_mu0 = p.tv_mu0 + (int64_t)f * dx;
      g_S0 = p.tv_Sigma0 + (int64_t)f * nxx;
    } else {
      g_mu0 = p.params + p.param_off[f];
      g_S0 = g_mu0 + dx;
    }
    const int64_t blk0 = TV ? (int64_t)f * p.tv_blocks : 0;
    load_params(blk0);
    // Q at this wave's tiles of Sigma_p, in registers (reloaded with a step's parameters).
    // Q is symmetric (a covariance): reading Q[qc][qr] makes the 16 lanes of a row read one
    // contiguous 128-byte run.
    float qv[TX][4];
    auto load_q = [&]() {
#pragma unroll
      for (int m = 0; m < TX; ++m) {
        const int r = jw + m * nj, qc = 16 * rb + li;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int qr = 16 * r + 4 * lk + e;
          qv[m][e] = (r < tn && qr < dx && qc < dx) ? (float)gQ_cur[qc + (int64_t)qr * dx] : 0.f;
        }
      }
    };
    load_q();
    // Sigma0 from its upper triangle, as Symmetric(:U)
    for (int e = tid; e < dxp * dxp; e += NT) {
      const int i = e % dxp, j = e / dxp;
      const int lo = i < j ? i : j, hi = i < j ? j : i;
      B0[swz<DXP>(i, j)] = (hi < dx) ? (float)g_S0[lo + (int64_t)hi * dx] : 0.f;
    }
    for (int e = tid; e < dxp; e += NT) mu[e] = e < dx ? (float)g_mu0[e] : 0.f;
    __syncthreads();
    const float jit = (float)p.jitter;
    double ll = 0.0;
    int ok = 1;
    const double* yrow = p.ys + p.y_off[f];
    const int Tn = p.Tf ? p.Tf[f] : p.T;
    float ynext = (tid < dy && Tn > 0) ? (float)yrow[tid] : 0.f;
    for (int t = 0; t < Tn; ++t) {
      // The lane and shape indices pass through empty asm each step, so the compiler
      // recomputes the step's LDS addresses instead of keeping them all live across the
      // loop (which took the kernel past its register budget).
      int li = li0, lk = lk0;
      asm volatile("" : "+v"(li), "+v"(lk));
      int dx = dx0, dy = dy0, dxp = dxp0, dyp = dyp0, tn = tn0, tny = tny0;
      asm volatile("" : "+s"(dx), "+s"(dy), "+s"(dxp), "+s"(dyp), "+s"(tn), "+s"(tny));
      if constexpr (TV) {
        if (p.tv_blocks > 1 && t > 0) {  // this step's parameters (the last step ended synced)
          load_params(blk0 + t);
          load_q();
          __syncthreads();
        }
      }
      // P1: T1 = A Sigma (Sigma symmetric, so its rows serve as Sigma^T's);
      // mu_hat = A mu + b; y_t (loaded one step ahead)
      for (int c = jw; c < tn; c += 2 * nj) {  // two tiles at a time, chains interleaved
        const int c1 = c + nj < tn ? c + nj : c;
        f4v bb[TX], bb1[TX];
#pragma unroll
        for (int kb = 0; kb < TX; ++kb)
          if (kb < tn) {
            bb[kb] = *reinterpret_cast<const f4v*>(B0 + swz<DXP>(16 * c + li, 16 * kb + 4 * lk));
            bb1[kb] = *reinterpret_cast<const f4v*>(B0 + swz<DXP>(16 * c1 + li, 16 * kb + 4 * lk));
          }
        f4v acc = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < TX; ++kb)
          if (kb < tn) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[kb][u], bb[kb][u], acc, 0, 0, 0);
              acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(areg[kb][u], bb1[kb][u], acc1, 0, 0, 0);
            }
          }
        kf_store<DXP>(B1, 16 * rb, 16 * c, acc, li, lk);
        if (c1 != c) kf_store<DXP>(B1, 16 * rb, 16 * c1, acc1, li, lk);
      }
      if (jw == 0) {
        f4v sp = {0.f, 0.f, 0.f, 0.f};  // four partial sums: short dependency chains
#pragma unroll
        for (int kb = 0; kb < TX; ++kb) {
          if (kb < tn) {
            const f4v m4 = *reinterpret_cast<const f4v*>(mu + 16 * kb + 4 * lk);
            sp += areg[kb] * m4;
          }
        }
        float s = (sp[0] + sp[1]) + (sp[2] + sp[3]);
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        const int i = 16 * rb + li;
        if (lk == 0) muh[i] = (i < dx) ? s + bv[i] : 0.f;
      }
      if (tid < dyp) {
        yv[tid] = ynext;
        if (t + 1 < Tn && tid < dy) ynext = (float)yrow[(size_t)(t + 1) * dy + tid];
      }
      __syncthreads();
      // P2: Sigma_p = T1 A' + Q, stored as Sigma_p^T over Sigma; z = y - H mu_hat - c
#pragma unroll
      for (int m = 0; m < TX; m += 2) {  // two tiles at a time, chains interleaved
        const int r = jw + m * nj, r1 = jw + (m + 1) * nj;
        if (r < tn) {
          const bool two = (m + 1 < TX) && r1 < tn;
          const int r1c = two ? r1 : r;
          f4v aa[TX], aa1[TX];
#pragma unroll
          for (int kb = 0; kb < TX; ++kb)
            if (kb < tn) {
              aa[kb] = *reinterpret_cast<const f4v*>(B1 + swz<DXP>(16 * r + li, 16 * kb + 4 * lk));
              aa1[kb] = *reinterpret_cast<const f4v*>(B1 + swz<DXP>(16 * r1c + li, 16 * kb + 4 * lk));
            }
          f4v acc = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kb = 0; kb < TX; ++kb)
            if (kb < tn) {
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(aa[kb][u], areg[kb][u], acc, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(aa1[kb][u], areg[kb][u], acc1, 0, 0, 0);
              }
            }
#pragma unroll
          for (int e = 0; e < 4; ++e) acc[e] += qv[m][e];
          kf_store_t<DXP>(B0, 16 * r, 16 * rb, acc, li, lk);
          if (m + 1 < TX && two) {
#pragma unroll
            for (int e = 0; e < 4; ++e) acc1[e] += qv[m + 1 < TX ? m + 1 : m][e];
            kf_store_t<DXP>(B0, 16 * r1, 16 * rb, acc1, li, lk);
          }
        }
      }
      for (int q = tid; q < dyp; q += NT) {
        f4v sp = {0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < dxp; k += 4) {
          const f4v h4 = *reinterpret_cast<const f4v*>(Hm + swz<DXP>(q, k));
          const f4v m4 = *reinterpret_cast<const f4v*>(muh + k);
          sp += h4 * m4;
        }
        zv[q] = (q < dy) ? yv[q] - (cv[q] + ((sp[0] + sp[1]) + (sp[2] + sp[3]))) : 0.f;
      }
      __syncthreads();
      // P3: P^T = Sigma_p^T H' (dxp x dyp), kept as is and transposed (P = H Sigma_p)
      kf_mm_tiles<DXP, DXP>(
          B0, Hm, dxp, wave, NW, tn * tny, li, lk,
          [&](int t, int& r0, int& c0) { r0 = 16 * (t / tny); c0 = 16 * (t % tny); },
          [&](int r0, int c0, f4v acc) {
            kf_store<DYP>(PT, r0, c0, acc, li, lk);
            kf_store_t<DXP>(Pm, r0, c0, acc, li, lk);
          });
      __syncthreads();
      // P4: S = P H' + R from its upper tiles, mirrored (the upper triangle is what the
      // factorisation of PDMat(S) reads)
      for (int tile = wave; tile < tny * (tny + 1) / 2; tile += NW) {
        const int r = tile == 2 ? 1 : 0, c = tile == 0 ? 0 : 1;  // (0,0), (0,1), (1,1)
        f4v acc = kf_mm<DXP, DXP>(Pm, Hm, 16 * r, 16 * c, dxp, li, lk);
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] += Rm[swz<DYP>(16 * r + 4 * lk + e, 16 * c + li)];
        if (r < c) {
          kf_store<DYP>(Sm, 16 * r, 16 * c, acc, li, lk);
          kf_store_t<DYP>(Sm, 16 * r, 16 * c, acc, li, lk);
        } else {
          const int j = 16 * c + li;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int ii = 16 * r + 4 * lk + e;
            if (ii <= j) {
              Sm[swz<DYP>(ii, j)] = acc[e];
              Sm[swz<DYP>(j, ii)] = acc[e];
            }
          }
        }
      }
      __syncthreads();
      // P5: S^-1 in FP64, log det S and z' S^-1 z on wave 0, which hands the inverse to the
      // other waves through LDS (B1 is free from P3 to P7)
      if (wave == 0) {
        SpdInv inv;
        double ldet, qd;
        const int good = spd_inverse_wave<DYP>(Sm, dy, zv, inv, &ldet, &qd);
        if (lane == 0) {
          iscal[1] = good;
          ll += -0.5 * ((double)dy * kLog2Pi + ldet + qd);
        }
        if (good) {
          if constexpr (TX == 1)
            gain_rows<DYP>(inv, dy, PT, Km, 0);  // the only row block
          else
            spd_inv_store<DYP == 32 ? 4 : 1>(inv, reinterpret_cast<double*>(B1));
        }
      }
      __syncthreads();
      if (!iscal[1]) {
        ok = 0;
        break;
      }
      // P6: K = P^T S^-1 in FP64 on the matrix cores, a row block a wave
      if constexpr (TX > 1) {
        if (wave < tn) {
          SpdInv inv;
          spd_inv_load<DYP == 32 ? 4 : 1>(inv, reinterpret_cast<const double*>(B1));
          for (int r = wave; r < tn; r += NW) gain_rows<DYP>(inv, dy, PT, Km, r);
        }
        __syncthreads();
      }
      // P7: T2 = Sigma_p (I - K H)' = Sigma_p - P^T K', stored as T2^T over T1;
      // mu = mu_hat + K z
      kf_mm_tiles<DYP, DYP>(
          PT, Km, dyp, wave, NW, tn * tn, li, lk,
          [&](int t, int& r0, int& c0) { r0 = 16 * (t / tn); c0 = 16 * (t % tn); },
          [&](int r0, int c0, f4v pk) {
            kf_store_t<DXP>(B1, r0, c0, kf_load_t<DXP>(B0, r0, c0, li, lk) - pk, li, lk);
          });
      for (int i = tid; i < dxp; i += NT) {
        f4v sp = {0.f, 0.f, 0.f, 0.f};
        for (int q = 0; q < dyp; q += 4) {
          const f4v k4 = *reinterpret_cast<const f4v*>(Km + swz<DYP>(i, q));
          const f4v z4 = *reinterpret_cast<const f4v*>(zv + q);
          sp += k4 * z4;
        }
        mu[i] = (i < dx) ? muh[i] + ((sp[0] + sp[1]) + (sp[2] + sp[3])) : 0.f;
      }
      __syncthreads();
      // P8: U = R K' - H T2 (dyp x dxp), stored as U^T over P^T
      for (int tile = wave; tile < tny * tn; tile += NW) {
        const int r = tile / tn, c = tile % tn;
        const f4v ht = kf_mm<DXP, DXP>(Hm, B1, 16 * r, 16 * c, dxp, li, lk);
        const f4v rk = kf_mm<DYP, DYP>(Rm, Km, 16 * r, 16 * c, dyp, li, lk);
        kf_store_t<DYP>(PT, 16 * r, 16 * c, rk - ht, li, lk);
      }
      __syncthreads();
      // P9: Sigma = Symmetric(:U)(T2 + K U) + jitter I: the upper tiles, mirrored
      kf_mm_tiles<DYP, DYP>(
          Km, PT, dyp, wave, NW, tn * (tn + 1) / 2, li, lk,
          [&](int t, int& r0, int& c0) {  // upper tile t, row by row
            int r = 0, c = t;
            while (c >= tn - r) {
              c -= tn - r;
              ++r;
            }
            r0 = 16 * r;
            c0 = 16 * (c + r);
          },
          [&](int r0, int c0, f4v ku) {
            const f4v v = kf_load_t<DXP>(B1, r0, c0, li, lk) + ku;
            if (r0 < c0) {
              kf_store<DXP>(B0, r0, c0, v, li, lk);
              kf_store_t<DXP>(B0, r0, c0, v, li, lk);
            } else {
              const int j = c0 + li;
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                const int i = r0 + 4 * lk + e;
                if (i <= j) {
                  const float w = v[e] + ((i == j && i < dx) ? jit : 0.f);
                  B0[swz<DXP>(i, j)] = w;
                  B0[swz<DXP>(j, i)] = w;
                }
              }
            }
          });
      __syncthreads();
    }
    if (tid == 0) {
      p.ll_out[f] = ok ? ll : NAN;
      p.status[f] = ok ? SSM_OK : SSM_ERR_NOT_PD;
    }
    if (p.mu_out && ok) {
      const int64_t mo = p.mu_off[f], co = p.cov_off[f];
      for (int i = tid; i < dx; i += NT) p.mu_out[mo + i] = mu[i];
      for (int e = tid; e < dx * dx; e += NT) {
        const int i = e % dx, j = e / dx;  // column-major output
        p.Sigma_out[co + e] = B0[swz<DXP>(i, j)];
      }
    }
    __syncthreads();
  }
}

// Filters with Dx <= 4 and Dy <= 4 in the kf_loglikelihood layout (the TV arrays), one
// per lane with every matrix in registers: the recursion of k_kf_batch (Cholesky of S with
// its log-determinant and the solves, Joseph form, Symmetric(:U) plus jitter), all in FP64.
// Such filters are too small for matrix-core tiles or a workgroup each, where a step costs
// about the same for Dx = 2 as for Dx = 16; at this size FP64 fits the registers (256 VGPRs,
// no spills) and leaves only the inputs' own conditioning in the error.  Matrices are
// zero-padded to M, so the fixed M-term sums equal the dx- and dy-term ones.  One instance,
// M = 4, serves every shape: at B = 65536 (one wave per SIMD) instances sized to the filter
// (M = 1..3) measured no faster, and at 2 x 2 slower (1.19 against 0.86 ms for T = 100), as
// did loading the next step's parameters one step ahead.
constexpr int kTinyD = 4;
constexpr int kTinyThreads = 64;

template <int M>
__global__ void __launch_bounds__(kTinyThreads) k_kf_tiny(KfBatchParams p) {
  const int f = blockIdx.x * kTinyThreads + threadIdx.x;
  if (f >= p.count) return;
  const int dx = p.Dx[f], dy = p.Dy[f];
  const int64_t nxx = (int64_t)dx * dx, nyx = (int64_t)dy * dx, nyy = (int64_t)dy * dy;
  double mu[M], S[M][M], A[M][M], Q[M][M], H[M][M], R[M][M], b[M], c[M];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    mu[i] = (i < dx) ? p.tv_mu0[kf_0at(p, f, dx, i)] : 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j)
      S[i][j] = (i < dx && j < dx) ? p.tv_Sigma0[kf_0at(p, f, nxx, i + j * dx)] : 0.0;
  }
  auto load = [&](int t) {
#pragma unroll
    for (int i = 0; i < M; ++i) {
      b[i] = (i < dx) ? p.tv_b[kf_at(p, f, t, dx, i)] : 0.0;
      c[i] = (i < dy) ? p.tv_c[kf_at(p, f, t, dy, i)] : 0.0;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        A[i][j] = (i < dx && j < dx) ? p.tv_A[kf_at(p, f, t, nxx, i + j * dx)] : 0.0;
        Q[i][j] = (i < dx && j < dx) ? p.tv_Q[kf_at(p, f, t, nxx, i + j * dx)] : 0.0;
        H[i][j] = (i < dy && j < dx) ? p.tv_H[kf_at(p, f, t, nyx, i + j * dy)] : 0.0;
        R[i][j] = (i < dy && j < dy) ? p.tv_R[kf_at(p, f, t, nyy, i + j * dy)] : 0.0;
      }
    }
  };
  load(0);
  const double jit = p.jitter;
  const int Tn = p.Tf ? p.Tf[f] : p.T;
  double ll = 0.0;
  int ok = 1;
  for (int t = 0; t < Tn; ++t) {
    if (p.tv_blocks > 1 && t > 0) load(t);
    double y[M];
#pragma unroll
    for (int q = 0; q < M; ++q) y[q] = (q < dy) ? p.ys[kf_yat(p, f, t, dy, q)] : 0.0;
    // predict: mu_hat = A mu + b, Sigma = A Sigma A' + Q
    double mh[M], T1[M][M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double s = b[i];
#pragma unroll
      for (int k = 0; k < M; ++k) s += A[i][k] * mu[k];
      mh[i] = s;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double u = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) u += A[i][k] * S[k][j];
        T1[i][j] = u;
      }
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double u = Q[i][j];
#pragma unroll
        for (int k = 0; k < M; ++k) u += T1[i][k] * A[j][k];
        S[i][j] = u;
      }
    // update: z = y - H mu_hat - c, P = H Sigma, S_y = P H' + R
    double z[M], P[M][M];
    double Sy[M][M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      double s = c[q];
#pragma unroll
      for (int k = 0; k < M; ++k) s += H[q][k] * mh[k];
      z[q] = (q < dy) ? y[q] - s : 0.0;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double u = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) u += H[q][k] * S[k][j];
        P[q][j] = u;
      }
    }
#pragma unroll
    for (int q = 0; q < M; ++q)
#pragma unroll
      for (int r = 0; r < M; ++r) {
        double u = R[q][r];
#pragma unroll
        for (int k = 0; k < M; ++k) u += P[q][k] * H[r][k];
        Sy[q][r] = (double)u;
      }
    // FP64 Cholesky of S_y (lower triangle read), log-determinant, L^-1 z
    double L[M][M], ldet = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j) {
#pragma unroll
      for (int i = 0; i < M; ++i) L[i][j] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < M; ++j) {
      if (j < dy) {
        double d = Sy[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        if (!(d > 0.0)) ok = 0;
        const double ljj = sqrt(d > 0.0 ? d : 1.0);
        L[j][j] = ljj;
        ldet += 2.0 * log(ljj);
#pragma unroll
        for (int i = j + 1; i < M; ++i) {
          if (i < dy) {
            double v = Sy[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
            L[i][j] = v / ljj;
          }
        }
      }
    }
    if (!ok) break;
    double w[M], qd = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      if (i < dy) {
        double v = (double)z[i];
#pragma unroll
        for (int k = 0; k < i; ++k) v -= L[i][k] * w[k];
        w[i] = v / L[i][i];
        qd += w[i] * w[i];
      } else {
        w[i] = 0.0;
      }
    }
    ll += -0.5 * ((double)dy * kLog2Pi + ldet + qd);
    // K' = S_y^-1 P (forward and back substitution per column of P), K = (K')'
    double K[M][M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      double v[M];
#pragma unroll
      for (int i = 0; i < M; ++i) {
        if (i < dy) {
          double s = (double)P[i][j];
#pragma unroll
          for (int k = 0; k < i; ++k) s -= L[i][k] * v[k];
          v[i] = s / L[i][i];
        } else {
          v[i] = 0.0;
        }
      }
#pragma unroll
      for (int i = M - 1; i >= 0; --i) {
        if (i < dy) {
          double s = v[i];
#pragma unroll
          for (int k = i + 1; k < M; ++k)
            if (k < dy) s -= L[k][i] * v[k];
          v[i] = s / L[i][i];
        }
      }
#pragma unroll
      for (int q = 0; q < M; ++q) K[j][q] = (q < dy && j < dx) ? v[q] : 0.0;
    }
    // mu = mu_hat + K z; Joseph: Sigma = (I - K H) Sigma (I - K H)' + K R K'
    double Mm[M][M], KR[M][M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double s = mh[i];
#pragma unroll
      for (int q = 0; q < M; ++q) s += K[i][q] * z[q];
      mu[i] = (i < dx) ? s : 0.0;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double u = (i == j && i < dx) ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < M; ++q) u -= K[i][q] * H[q][j];
        Mm[i][j] = u;
        double r = 0.0;
#pragma unroll
        for (int q = 0; q < M; ++q) r += K[i][q] * R[q][j];
        KR[i][j] = r;
      }
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double u = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) u += Mm[i][k] * S[k][j];
        T1[i][j] = u;
      }
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
      for (int j = i; j < M; ++j) {  // the upper triangle, mirrored below (Symmetric(:U))
        double u = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) u += T1[i][k] * Mm[j][k];
#pragma unroll
        for (int k = 0; k < M; ++k) u += KR[i][k] * K[j][k];
        S[i][j] = u;
      }
#pragma unroll
    for (int i = 0; i < M; ++i) {
#pragma unroll
      for (int j = 0; j < i; ++j) S[i][j] = S[j][i];
      if (i < dx) S[i][i] += jit;
    }
  }
  p.ll_out[f] = ok ? ll : NAN;
  p.status[f] = ok ? SSM_OK : SSM_ERR_NOT_PD;
}


// ---- gradients of kf_loglikelihood for filters up to 4 x 4 -----------------------------
// The reverse rule of GF/src/integrations/kalman_rrule.jl:22-107 with the backward recursion
// of GF/src/algorithms/kalman_gradient.jl:103-263 (Parellier et al.), one filter per lane in
// FP64, zero-padded to N = max(Dx, Dy) as k_kf_tiny.  The forward pass is k_kf_tiny's and
// leaves a cache per step in a filter-minor array (slot e of step t of filter f at
// (t kC + e) B + f with kC = 4 N + 4 N^2 slots a step, so the lanes of a wave touch
// consecutive words; only the slots inside dx and dy are written and read back); the
// backward pass walks the steps in reverse in the same lane.
constexpr int kGradCache = 80;  // slots a step at N = 4: mu_prev, S_prev, mu_pred, S_pred, mu_f, z, S, K

namespace {
template <int N>
__device__ __forceinline__ void mm4(const double (*a)[N], const double (*b)[N],
                                    double (*c)[N]) {  // c = a b
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[i][k] * b[k][j];
      c[i][j] = s;
    }
}
template <int N>
__device__ __forceinline__ void mtm4(const double (*a)[N], const double (*b)[N],
                                     double (*c)[N]) {  // c = a' b
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[k][i] * b[k][j];
      c[i][j] = s;
    }
}
// N x N helpers.  Lower Cholesky factor of the leading n x n block of a (identity past
// n); false if not PD.
template <int N>
__device__ __forceinline__ bool chol4(const double (*a)[N], int n, double (*L)[N]) {
  bool ok = true;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) L[i][j] = (i == j && i >= n) ? 1.0 : 0.0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    if (j < n) {
      double d = a[j][j];
#pragma unroll
      for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
      if (!(d > 0.0)) ok = false;
      L[j][j] = sqrt(d > 0.0 ? d : 1.0);
#pragma unroll
      for (int i = j + 1; i < N; ++i) {
        if (i < n) {
          double v = a[i][j];
#pragma unroll
          for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
          L[i][j] = v / L[j][j];
        }
      }
    }
  }
  return ok;
}
// x = (L L')^-1 v
template <int N>
__device__ __forceinline__ void cholsolve4(const double (*L)[N], const double* v, double* x) {
  double y[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double s = v[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
    y[i] = s / L[i][i];
  }
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int k = i + 1; k < N; ++k) s -= L[k][i] * x[k];
    x[i] = s / L[i][i];
  }
}
}  // namespace

template <int N>
__global__ void __launch_bounds__(kTinyThreads) k_kf_tiny_grad(KfBatchParams p, KfGradOut g) {
  using MN = double[N][N];
  // cache slots of a step: mu_prev, S_prev, mu_pred, S_pred, mu_f, z, S, K
  constexpr int oMu = 0, oS = N, oMp = N + N * N, oSp = 2 * N + N * N, oMf = 2 * N + 2 * N * N,
                oZ = 3 * N + 2 * N * N, oSy = 4 * N + 2 * N * N, oK = 4 * N + 3 * N * N,
                kC = 4 * N + 4 * N * N;
  static_assert(kC <= kGradCache, "cache slots");
  const int f = blockIdx.x * kTinyThreads + threadIdx.x;
  if (f >= p.count) return;
  const int dx = p.Dx[f], dy = p.Dy[f];
  const int64_t nxx = (int64_t)dx * dx, nyx = (int64_t)dy * dx, nyy = (int64_t)dy * dy;
  const int B = g.B;
  auto C = [&](int t, int e) -> double& { return g.cache[((size_t)t * kC + e) * B + f]; };
  double mu[N], S[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    mu[i] = (i < dx) ? p.tv_mu0[kf_0at(p, f, dx, i)] : 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j)
      S[i][j] = (i < dx && j < dx) ? p.tv_Sigma0[kf_0at(p, f, nxx, i + j * dx)] : 0.0;
  }
  auto blk = [&](int t) -> int64_t { return (int64_t)f * p.tv_blocks + (p.tv_blocks > 1 ? t : 0); };
  auto loadM = [&](const double* base, int t, int64_t E, int r, int c, MN X) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) X[i][j] = (i < r && j < c) ? base[kf_at(p, f, t, E, i + j * r)] : 0.0;
  };
  auto loadV = [&](const double* base, int t, int n, double* v) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = (i < n) ? base[kf_at(p, f, t, n, i)] : 0.0;
  };
  const int Tn = p.T;
  double ll = 0.0;
  int ok = 1;
  // ---- forward (kalman_rrule.jl:27-50), caching each step ----
  for (int t = 0; t < Tn && ok; ++t) {
    MN A, Q, H, R;
    double b[N], c[N];
    loadM(p.tv_A, t, nxx, dx, dx, A);
    loadM(p.tv_Q, t, nxx, dx, dx, Q);
    loadM(p.tv_H, t, nyx, dy, dx, H);
    loadM(p.tv_R, t, nyy, dy, dy, R);
    loadV(p.tv_b, t, dx, b);
    loadV(p.tv_c, t, dy, c);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i < dx) C(t, oMu + i) = mu[i];
#pragma unroll
      for (int j = 0; j < N; ++j)
        if (i < dx && j < dx) C(t, oS + N * i + j) = S[i][j];
    }
    double mp[N];
    MN T1, Sp;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = b[i];
#pragma unroll
      for (int kk = 0; kk < N; ++kk) s += A[i][kk] * mu[kk];
      mp[i] = s;
    }
    mm4<N>(A, S, T1);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double s = Q[i][j];
#pragma unroll
        for (int kk = 0; kk < N; ++kk) s += T1[i][kk] * A[j][kk];
        Sp[i][j] = s;
      }
    double z[N];
    MN P, Sy;
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double s = c[q];
#pragma unroll
      for (int kk = 0; kk < N; ++kk) s += H[q][kk] * mp[kk];
      z[q] = (q < dy) ? p.ys[kf_yat(p, f, t, dy, q)] - s : 0.0;
    }
    mm4<N>(H, Sp, P);
#pragma unroll
    for (int q = 0; q < N; ++q)
#pragma unroll
      for (int r = 0; r < N; ++r) {
        double s = R[q][r];
#pragma unroll
        for (int kk = 0; kk < N; ++kk) s += P[q][kk] * H[r][kk];
        Sy[q][r] = s;
      }
    MN L;
    if (!chol4<N>(Sy, dy, L)) {
      ok = 0;
      break;
    }
    double w[N], qd = 0.0, ldet = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = z[i];
#pragma unroll
      for (int kk = 0; kk < i; ++kk) s -= L[i][kk] * w[kk];
      w[i] = s / L[i][i];
      if (i < dy) {
        qd += w[i] * w[i];
        ldet += 2.0 * log(L[i][i]);
      }
    }
    ll += -0.5 * ((double)dy * kLog2Pi + ldet + qd);
    MN K;  // K = (S^-1 P)'
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double col[N], x[N];
#pragma unroll
      for (int q = 0; q < N; ++q) col[q] = P[q][j];
      cholsolve4<N>(L, col, x);
#pragma unroll
      for (int q = 0; q < N; ++q) K[j][q] = (q < dy && j < dx) ? x[q] : 0.0;
    }
    MN IKH, KR;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double u = (i == j && i < dx) ? 1.0 : 0.0, r = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) {
          u -= K[i][q] * H[q][j];
          r += K[i][q] * R[q][j];
        }
        IKH[i][j] = u;
        KR[i][j] = r;
      }
    double mf[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = mp[i];
#pragma unroll
      for (int q = 0; q < N; ++q) s += K[i][q] * z[q];
      mf[i] = (i < dx) ? s : 0.0;
    }
    mm4<N>(IKH, Sp, T1);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i; j < N; ++j) {  // upper triangle, mirrored (Symmetric(:U)), plus jitter
        double s = 0.0;
#pragma unroll
        for (int kk = 0; kk < N; ++kk) s += T1[i][kk] * IKH[j][kk] + KR[i][kk] * K[j][kk];
        S[i][j] = s;
      }
#pragma unroll
    for (int i = 0; i < N; ++i) {
#pragma unroll
      for (int j = 0; j < i; ++j) S[i][j] = S[j][i];
      if (i < dx) S[i][i] += p.jitter;
      mu[i] = mf[i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i < dx) {
        C(t, oMp + i) = mp[i];
        C(t, oMf + i) = mf[i];
      }
      if (i < dy) C(t, oZ + i) = z[i];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        if (i < dx && j < dx) C(t, oSp + N * i + j) = Sp[i][j];
        if (i < dy && j < dy) C(t, oSy + N * i + j) = Sy[i][j];
        if (i < dx && j < dy) C(t, oK + N * i + j) = K[i][j];
      }
    }
  }
  p.ll_out[f] = ok ? ll : NAN;
  p.status[f] = ok ? SSM_OK : SSM_ERR_NOT_PD;
  if (!ok) return;
  // ---- backward (kalman_rrule.jl:63-103), NLL convention; outputs are negated ----
  const bool acc = p.tv_blocks <= 1;  // one parameter set for all steps: sum the steps
  double dmu[N] = {};
  MN dS;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) dS[i][j] = 0.0;
  auto put = [&](double* out, int64_t off, int r, int cc, const MN X) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j)
        if (i < r && j < cc) {
          double& o = out[off + i + j * r];
          o = (acc ? o : 0.0) - X[i][j];
        }
  };
  auto putv = [&](double* out, int64_t off, int n, const double* v) {
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (i < n) {
        double& o = out[off + i];
        o = (acc ? o : 0.0) - v[i];
      }
  };
  if (acc) {  // the sums start at zero
    const int64_t k = blk(0);
#pragma unroll
    for (int i = 0; i < N * N; ++i) {
      if (i < nxx) {
        g.A[k * nxx + i] = 0.0;
        g.Q[k * nxx + i] = 0.0;
      }
      if (i < nyx) g.H[k * nyx + i] = 0.0;
      if (i < nyy) g.R[k * nyy + i] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i < dx) g.b[k * dx + i] = 0.0;
      if (i < dy) g.c[k * dy + i] = 0.0;
    }
  }
  for (int t = Tn - 1; t >= 0; --t) {
    const int64_t k = blk(t);
    MN A, H, R, Sp, Sy, K, Sprev;
    double mprev[N], mp[N], mf[N], z[N];
    loadM(p.tv_A, t, nxx, dx, dx, A);
    loadM(p.tv_H, t, nyx, dy, dx, H);
    loadM(p.tv_R, t, nyy, dy, dy, R);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      // the padding outside dx and dy was zero when cached
      mprev[i] = (i < dx) ? C(t, oMu + i) : 0.0;
      mp[i] = (i < dx) ? C(t, oMp + i) : 0.0;
      mf[i] = (i < dx) ? C(t, oMf + i) : 0.0;
      z[i] = (i < dy) ? C(t, oZ + i) : 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        Sprev[i][j] = (i < dx && j < dx) ? C(t, oS + N * i + j) : 0.0;
        Sp[i][j] = (i < dx && j < dx) ? C(t, oSp + N * i + j) : 0.0;
        Sy[i][j] = (i < dy && j < dy) ? C(t, oSy + N * i + j) : 0.0;
        K[i][j] = (i < dx && j < dy) ? C(t, oK + N * i + j) : 0.0;
      }
    }
    MN L, Sinv, IKH;
    chol4<N>(Sy, dy, L);
    double Siz[N];
    cholsolve4<N>(L, z, Siz);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double e[N], x[N];
#pragma unroll
      for (int i = 0; i < N; ++i) e[i] = (i == j) ? 1.0 : 0.0;
      cholsolve4<N>(L, e, x);
#pragma unroll
      for (int i = 0; i < N; ++i) Sinv[i][j] = (i < dy && j < dy) ? x[i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double u = (i == j && i < dx) ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) u -= K[i][q] * H[q][j];
        IKH[i][j] = u;
      }
    double Ktd[N];  // K' dmu
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < N; ++i) s += K[i][q] * dmu[i];
      Ktd[q] = s;
    }
    // gradient_c (kalman_gradient.jl:256-263)
    {
      double gc[N];
#pragma unroll
      for (int q = 0; q < N; ++q) gc[q] = -Siz[q] - Ktd[q];
      putv(g.c, k * dy, dy, gc);
    }
    // gradient_H (kalman_gradient.jl:216-246)
    {
      MN Pf, HSp, T2, gH;
      mm4<N>(IKH, Sp, Pf);
      mm4<N>(H, Sp, HSp);
      MN KtdS;  // K' dS
      mtm4<N>(K, dS, KtdS);
      mm4<N>(KtdS, Pf, T2);
      double SizTHSp[N], dmuPf[N];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        double s = 0.0, u = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) s += Siz[q] * HSp[q][j];
#pragma unroll
        for (int i = 0; i < N; ++i) u += dmu[i] * Pf[i][j];
        SizTHSp[j] = s;
        dmuPf[j] = u;
      }
#pragma unroll
      for (int q = 0; q < N; ++q)
#pragma unroll
        for (int j = 0; j < N; ++j) {
          double s = 0.0;
#pragma unroll
          for (int r = 0; r < N; ++r) s += Sinv[q][r] * HSp[r][j];
          gH[q][j] = s - Siz[q] * mp[j] - Siz[q] * SizTHSp[j] + Siz[q] * dmuPf[j] -
                     Ktd[q] * mf[j] - 2.0 * T2[q][j];
        }
      put(g.H, k * nyx, dy, dx, gH);
    }
    // gradient_R (kalman_gradient.jl:160-174)
    {
      MN KtdSK, gR;
      MN KtdS;
      mtm4<N>(K, dS, KtdS);
      mm4<N>(KtdS, K, KtdSK);
#pragma unroll
      for (int q = 0; q < N; ++q)
#pragma unroll
        for (int r = 0; r < N; ++r)
          gR[q][r] = KtdSK[q][r] - 0.5 * (Ktd[q] * Siz[r] + Siz[q] * Ktd[r]) +
                     0.5 * (Sinv[q][r] - Siz[q] * Siz[r]);
      put(g.R, k * nyy, dy, dy, gR);
    }
    // backward_gradient_update (kalman_gradient.jl:103-124)
    double dmp[N];
    MN dSp;
    {
      double HtSiz[N], Riz[N], HtRiz[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double s = 0.0, u = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) {
          s += H[q][i] * Siz[q];
          u += IKH[q][i] * dmu[q];
        }
        HtSiz[i] = s;
        dmp[i] = u - s;
      }
      MN LR;
      chol4<N>(R, dy, LR);
      cholsolve4<N>(LR, z, Riz);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) s += H[q][i] * (q < dy ? Riz[q] : 0.0);
        HtRiz[i] = s;
      }
      MN inner, T3, HtSinvH;
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
          inner[i][j] = dS[i][j] + 0.5 * (dmu[i] * HtRiz[j] + HtRiz[i] * dmu[j]);
      mtm4<N>(IKH, inner, T3);   // IKH' inner
      mm4<N>(T3, IKH, dSp);      // IKH' inner IKH
      MN SinvH;
      mm4<N>(Sinv, H, SinvH);
      mtm4<N>(H, SinvH, HtSinvH);
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
          dSp[i][j] += 0.5 * (HtSinvH[i][j] - HtSiz[i] * HtSiz[j]);
    }
    // dynamics parameters (kalman_gradient.jl:150-152, :196-208)
    putv(g.b, k * dx, dx, dmp);
    put(g.Q, k * nxx, dx, dx, dSp);
    {
      MN T4, T5, gA;
      mm4<N>(dSp, A, T4);
      mm4<N>(T4, Sprev, T5);
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) gA[i][j] = dmp[i] * mprev[j] + 2.0 * T5[i][j];
      put(g.A, k * nxx, dx, dx, gA);
    }
    // backward_gradient_predict (kalman_gradient.jl:139-143)
    {
      MN T6;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < N; ++q) s += A[q][i] * dmp[q];
        dmu[i] = s;
      }
      mtm4<N>(A, dSp, T6);
      mm4<N>(T6, A, dS);
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i < dx) g.mu0[(int64_t)f * dx + i] = -dmu[i];
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (i < dx && j < dx) g.Sigma0[(int64_t)f * nxx + i + j * dx] = -dS[i][j];
  }
}

// ---- square-root Kalman filter for filters up to 4 x 4 -------------------------------
// GF/src/algorithms/srkf.jl:93-176, one filter per lane in FP64, matrices padded to N.  The
// covariance is carried as its upper Cholesky factor U (Sigma = U' U).  Predict: the R of
// a QR factorisation of [U_Q; U A'] (2N x N), rows made to have a positive diagonal.
// Update: the R of [[U_R, 0], [U H', U]] (2N x 2N), sign-corrected, gives U_S (the factor
// of S), X = U_S'^-1 H P and the filtered factor; w = U_S'^-1 z, mu += X' w and
// ll += -(w'w + 2 sum log diag U_S + Dy log 2 pi) / 2.  Jitter is added to the factor's
// diagonal, as srkf_update does.  The QR is by Givens rotations below the diagonal, column
// by column; padding rows and columns stay zero through every rotation, so the blocks sit
// at fixed offsets (N) whatever Dx and Dy.
namespace {
template <int M, int C>
__device__ __forceinline__ void givens_r(double (&X)[M][C], int ncols) {
#pragma unroll
  for (int j = 0; j < C; ++j) {
    if (j >= ncols) break;
#pragma unroll
    for (int i = j + 1; i < M; ++i) {
      const double a = X[j][j], bb = X[i][j];
      if (bb != 0.0) {
        const double r = sqrt(a * a + bb * bb);
        const double cs = a / r, sn = bb / r;
#pragma unroll
        for (int k = 0; k < C; ++k) {
          if (k >= j) {
            const double u = X[j][k], v = X[i][k];
            X[j][k] = cs * u + sn * v;
            X[i][k] = cs * v - sn * u;
          }
        }
      }
    }
  }
  // _correct_cholesky_sign: rows scaled by the sign of their diagonal entry (0 stays 0)
#pragma unroll
  for (int i = 0; i < M && i < C; ++i) {
    const double d = X[i][i];
    const double sg = d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : 0.0);
#pragma unroll
    for (int k = 0; k < C; ++k) X[i][k] *= sg;
  }
}
}  // namespace

template <int N>
__global__ void __launch_bounds__(kTinyThreads) k_srkf_tiny(KfBatchParams p, double* mu_out,
                                                            double* U_out) {
  const int f = blockIdx.x * kTinyThreads + threadIdx.x;
  if (f >= p.count) return;
  const int dx = p.Dx[f], dy = p.Dy[f];
  const int64_t nxx = (int64_t)dx * dx, nyx = (int64_t)dy * dx, nyy = (int64_t)dy * dy;
  double mu[N], U[N][N];
  int ok = 1;
  {  // U = cholesky(Sigma0).U
    double S0[N][N], L[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      mu[i] = (i < dx) ? p.tv_mu0[kf_0at(p, f, dx, i)] : 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j)
        S0[i][j] = (i < dx && j < dx) ? p.tv_Sigma0[kf_0at(p, f, nxx, i + j * dx)] : 0.0;
    }
    ok &= chol4<N>(S0, dx, L);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) U[i][j] = (i < dx && j < dx) ? L[j][i] : 0.0;
  }
  auto loadM = [&](const double* base, int t, int64_t E, int r, int c, double (&X)[N][N]) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j)
        X[i][j] = (i < r && j < c) ? base[kf_at(p, f, t, E, i + j * r)] : 0.0;
  };
  double ll = 0.0;
  for (int t = 0; t < p.T && ok; ++t) {
    double A[N][N], Q[N][N], H[N][N], R[N][N], L[N][N];
    loadM(p.tv_A, t, nxx, dx, dx, A);
    loadM(p.tv_Q, t, nxx, dx, dx, Q);
    loadM(p.tv_H, t, nyx, dy, dx, H);
    loadM(p.tv_R, t, nyy, dy, dy, R);
    // predict (srkf_predict): mu = A mu + b; U = R of qr([U_Q; U A'])
    {
      double m2[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double s = (i < dx) ? p.tv_b[kf_at(p, f, t, dx, i)] : 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) s += A[i][j] * mu[j];
        m2[i] = s;
      }
#pragma unroll
      for (int i = 0; i < N; ++i) mu[i] = m2[i];
      ok &= chol4<N>(Q, dx, L);
      double Mp[2 * N][N];
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
          Mp[i][j] = (i < dx && j < dx) ? L[j][i] : 0.0;  // U_Q
          double s = 0.0;                                   // (U A')[i][j]
#pragma unroll
          for (int q = 0; q < N; ++q) s += U[i][q] * A[j][q];
          Mp[N + i][j] = s;
        }
      givens_r<2 * N, N>(Mp, dx);
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) U[i][j] = (i < dx && j < dx) ? Mp[i][j] : 0.0;
    }
    // update (srkf_update): R of qr([[U_R, 0], [U H', U]])
    ok &= chol4<N>(R, dy, L);
    double Mu[2 * N][2 * N];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        Mu[i][j] = (i < dy && j < dy) ? L[j][i] : 0.0;  // U_R
        Mu[i][N + j] = 0.0;
        double s = 0.0;                                   // (U H')[i][j]
#pragma unroll
        for (int q = 0; q < N; ++q) s += U[i][q] * H[j][q];
        Mu[N + i][j] = s;
        Mu[N + i][N + j] = U[i][j];
      }
    givens_r<2 * N, 2 * N>(Mu, 2 * N);
    double z[N], w[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double s = (q < dy) ? p.tv_c[kf_at(p, f, t, dy, q)] : 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s += H[q][j] * mu[j];
      z[q] = (q < dy) ? p.ys[kf_yat(p, f, t, dy, q)] - s : 0.0;
    }
    double qd = 0.0, ldet = 0.0;
#pragma unroll
    for (int q = 0; q < N; ++q) {  // w = U_S' \ z (U_S' lower triangular)
      if (q < dy) {
        double s = z[q];
#pragma unroll
        for (int r = 0; r < q; ++r) s -= Mu[r][q] * w[r];
        const double d = Mu[q][q];
        if (!(d > 0.0)) ok = 0;
        w[q] = s / d;
        qd += w[q] * w[q];
        ldet += 2.0 * log(d > 0.0 ? d : 1.0);
      } else {
        w[q] = 0.0;
      }
    }
    ll += -0.5 * (qd + ldet + (double)dy * kLog2Pi);
#pragma unroll
    for (int i = 0; i < N; ++i) {  // mu += (P H' U_S^-1) w = X' w
      double s = mu[i];
#pragma unroll
      for (int q = 0; q < N; ++q) s += Mu[q][N + i] * w[q];
      mu[i] = (i < dx) ? s : 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j)
        U[i][j] = (i < dx && j < dx) ? Mu[N + i][N + j] + ((i == j) ? p.jitter : 0.0) : 0.0;
    }
  }
  p.ll_out[f] = ok ? ll : NAN;
  p.status[f] = ok ? SSM_OK : SSM_ERR_NOT_PD;
  if (!ok) return;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i < dx) mu_out[(int64_t)f * dx + i] = mu[i];
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (i < dx && j < dx) U_out[(int64_t)f * nxx + i + j * dx] = U[i][j];
  }
}

}  // namespace ssm
using namespace ssm;

namespace {

constexpr int kClasses = 3;
struct ClassCfg {
  int dx_max, dy_max;
};
constexpr ClassCfg kCfg[kClasses] = {{16, 16}, {32, 16}, {64, 32}};

__global__ void k_zero_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

template <int DXP, int DYP, int NW>
int launch_class(ssm_ctx* ctx, hipStream_t s, const KfBatchParams& p, int grid_cap) {
  const size_t lds = KfLds<DXP, DYP>::kBytes;
  auto fn = p.tv_A ? k_kf_batch<DXP, DYP, NW, true> : k_kf_batch<DXP, DYP, NW, false>;
  SSM_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)fn,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  SSM_HIP_CHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * NW, lds));
  if (per_cu < 1) per_cu = 1;
  int grid = per_cu * (ctx->cu_count > 0 ? ctx->cu_count : 256);
  if (grid_cap > 0) grid = std::min(grid, grid_cap);
  grid = std::max(1, std::min(grid, p.count));
  hipLaunchKernelGGL(fn, dim3(grid), dim3(64 * NW), lds, s, p);
  SSM_HIP_CHECK(ctx, hipGetLastError());
  return SSM_OK;
}

// Waves per workgroup of each size class: candidates the autotuner times, the default
// before tuning, and the process-wide tuned choice (0 = not tuned).
// (0: no candidate; 16 waves of the large class would leave 128 VGPRs a lane)
constexpr int kNwCand[kClasses][3] = {{1, 2, 4}, {2, 4, 8}, {4, 8, 0}};
constexpr int kNwDefault[kClasses] = {1, 2, 4};
std::atomic<int> g_nw_tuned[kClasses] = {{0}, {0}, {0}};

int class_nw(int c) {
  const int v = g_nw_tuned[c].load();
  return v ? v : kNwDefault[c];
}

int launch_class_nw(ssm_ctx* ctx, hipStream_t s, const KfBatchParams& p, int c, int nw) {
  switch (c * 100 + nw) {
    case 1: return launch_class<16, 16, 1>(ctx, s, p, 0);
    case 2: return launch_class<16, 16, 2>(ctx, s, p, 0);
    case 4: return launch_class<16, 16, 4>(ctx, s, p, 0);
    case 102: return launch_class<32, 16, 2>(ctx, s, p, 0);
    case 104: return launch_class<32, 16, 4>(ctx, s, p, 0);
    case 108: return launch_class<32, 16, 8>(ctx, s, p, 0);
    case 204: return launch_class<64, 32, 4>(ctx, s, p, 0);
    case 208: return launch_class<64, 32, 8>(ctx, s, p, 0);
    default: break;
  }
  return set_error(ctx, SSM_ERR_ARG, "no Kalman kernel for class %d with %d waves", c, nw);
}

}  // namespace

namespace ssm {

int kf_group_plan(ssm_ctx* ctx, int B, const int32_t* Dx, const int32_t* Dy, const int32_t* Tf,
                  int32_t* order, int count[3], int start[3]) {
  std::vector<std::vector<int32_t>> cls(kClasses);
  for (int f = 0; f < B; ++f) {
    int c = -1;
    for (int k = 0; k < kClasses && c < 0; ++k)
      if (Dx[f] >= 1 && Dy[f] >= 1 && Dx[f] <= kCfg[k].dx_max && Dy[f] <= kCfg[k].dy_max) c = k;
    if (c < 0)
      return set_error(ctx, SSM_ERR_ARG, "filter %d: dims Dx=%d Dy=%d outside 1..64 / 1..32", f,
                       Dx[f], Dy[f]);
    cls[c].push_back(f);
  }
  int n = 0;
  for (int c = 0; c < kClasses; ++c) {
    std::stable_sort(cls[c].begin(), cls[c].end(), [&](int a, int b) {
      const double ca = ((double)Dx[a] * Dx[a] * Dx[a] + (double)Dy[a] * Dy[a] * Dy[a]) * Tf[a];
      const double cb = ((double)Dx[b] * Dx[b] * Dx[b] + (double)Dy[b] * Dy[b] * Dy[b]) * Tf[b];
      return ca > cb;
    });
    start[c] = n;
    count[c] = (int)cls[c].size();
    for (int f : cls[c]) order[n++] = f;
  }
  return SSM_OK;
}

int kf_group_launch(ssm_ctx* ctx, hipStream_t s, const KfGroupDev& g) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, g.counter, 64);
  for (int c = kClasses - 1; c >= 0; --c) {
    if (g.count[c] == 0) continue;
    KfBatchParams p;
    p.T = 0;
    p.Tf = g.Tf;
    p.count = g.count[c];
    p.order = g.order + g.start[c];
    p.counter = g.counter + 16 * c;
    p.Dx = g.Dx;
    p.Dy = g.Dy;
    p.params = g.params;
    p.param_off = g.param_off;
    p.ys = g.ys;
    p.y_off = g.y_off;
    p.ll_out = g.ll;
    p.status = g.status;
    p.mu_out = nullptr;
    p.Sigma_out = nullptr;
    p.mu_off = nullptr;
    p.cov_off = nullptr;
    const int rc = launch_class_nw(ctx, s, p, c, class_nw(c));
    if (rc) return rc;
  }
  return SSM_OK;
}

}  // namespace ssm

struct ssm_kf_batch {
  ssm_ctx* ctx = nullptr;
  int B = 0, T = 0;
  int32_t* dDx = nullptr;
  int32_t* dDy = nullptr;
  double* dparams = nullptr;
  int64_t* dpoff = nullptr;
  double* dys = nullptr;
  int64_t* dyoff = nullptr;
  double* dll = nullptr;
  int32_t* dstatus = nullptr;
  double* dmu = nullptr;
  double* dSig = nullptr;
  int64_t* dmuoff = nullptr;
  int64_t* dcovoff = nullptr;
  int32_t* dorder = nullptr;
  uint32_t* dcounter = nullptr;
  int count[kClasses] = {0, 0, 0};
  int start[kClasses] = {0, 0, 0};
  int64_t mu_len = 0, cov_len = 0;
  hipStream_t streams[kClasses] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[kClasses] = {nullptr, nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
};

static void kf_free(ssm_kf_batch* kb) {
  void* ptrs[] = {kb->dDx,    kb->dDy,     kb->dparams, kb->dpoff,  kb->dys,
                  kb->dyoff,  kb->dll,     kb->dstatus, kb->dmu,    kb->dSig,
                  kb->dmuoff, kb->dcovoff, kb->dorder,  kb->dcounter};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int c = 0; c < kClasses; ++c) {
    if (kb->streams[c]) (void)hipStreamDestroy(kb->streams[c]);
    if (kb->join[c]) (void)hipEventDestroy(kb->join[c]);
  }
  if (kb->fork) (void)hipEventDestroy(kb->fork);
  if (kb->ev0) (void)hipEventDestroy(kb->ev0);
  if (kb->ev1) (void)hipEventDestroy(kb->ev1);
}

template <typename X>
static bool dup(X** d, const X* h, size_t n) {
  if (hipMalloc(d, sizeof(X) * (n ? n : 1)) != hipSuccess) return false;
  return n == 0 || hipMemcpy(*d, h, sizeof(X) * n, hipMemcpyHostToDevice) == hipSuccess;
}

extern "C" int32_t ssm_kf_batch_create(ssm_ctx* ctx, int32_t B, const int32_t* Dx,
                                       const int32_t* Dy, const double* params,
                                       const int64_t* param_off, int32_t T, const double* ys,
                                       const int64_t* y_off, ssm_kf_batch** out) {
  if (!ctx || !out) return SSM_ERR_ARG;
  *out = nullptr;
  if (B < 1 || T < 1 || !Dx || !Dy || !params || !param_off || !ys || !y_off)
    return set_error(ctx, SSM_ERR_ARG, "ssm_kf_batch_create: bad arguments");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int64_t plen = 0, ylen = 0;
  std::vector<int64_t> muoff(B), covoff(B);
  std::vector<std::vector<int32_t>> cls(kClasses);
  int64_t mu_len = 0, cov_len = 0;
  for (int f = 0; f < B; ++f) {
    const int dx = Dx[f], dy = Dy[f];
    int c = -1;
    for (int k = 0; k < kClasses && c < 0; ++k)
      if (dx >= 1 && dy >= 1 && dx <= kCfg[k].dx_max && dy <= kCfg[k].dy_max) c = k;
    if (c < 0)
      return set_error(ctx, SSM_ERR_ARG,
                       "filter %d: dims Dx=%d Dy=%d outside 1..64 / 1..32", f, dx, dy);
    if (param_off[f] < 0 || y_off[f] < 0)
      return set_error(ctx, SSM_ERR_ARG, "filter %d: negative offset", f);
    const int64_t psz = 3LL * dx * dx + 2LL * dx + (int64_t)dy * dx + dy + (int64_t)dy * dy;
    plen = std::max(plen, param_off[f] + psz);
    ylen = std::max(ylen, y_off[f] + (int64_t)T * dy);
    cls[c].push_back(f);
    muoff[f] = mu_len;
    covoff[f] = cov_len;
    mu_len += dx;
    cov_len += (int64_t)dx * dx;
  }
  // costliest first within each class (largest-processing-time order for the queue)
  std::vector<int32_t> order;
  ssm_kf_batch* kb = new ssm_kf_batch();
  kb->ctx = ctx;
  kb->B = B;
  kb->T = T;
  for (int c = 0; c < kClasses; ++c) {
    std::stable_sort(cls[c].begin(), cls[c].end(), [&](int a, int b) {
      const double ca = (double)Dx[a] * Dx[a] * Dx[a] + (double)Dy[a] * Dy[a] * Dy[a];
      const double cb = (double)Dx[b] * Dx[b] * Dx[b] + (double)Dy[b] * Dy[b] * Dy[b];
      return ca > cb;
    });
    kb->start[c] = (int)order.size();
    kb->count[c] = (int)cls[c].size();
    order.insert(order.end(), cls[c].begin(), cls[c].end());
  }
  kb->mu_len = mu_len;
  kb->cov_len = cov_len;
  bool ok = dup(&kb->dDx, Dx, B) && dup(&kb->dDy, Dy, B) && dup(&kb->dparams, params, plen) &&
            dup(&kb->dpoff, param_off, B) && dup(&kb->dys, ys, ylen) &&
            dup(&kb->dyoff, y_off, B) && dup(&kb->dmuoff, muoff.data(), B) &&
            dup(&kb->dcovoff, covoff.data(), B) && dup(&kb->dorder, order.data(), B) &&
            hipMalloc(&kb->dll, sizeof(double) * B) == hipSuccess &&
            hipMalloc(&kb->dstatus, sizeof(int32_t) * B) == hipSuccess &&
            hipMalloc(&kb->dmu, sizeof(double) * std::max<int64_t>(1, mu_len)) == hipSuccess &&
            hipMalloc(&kb->dSig, sizeof(double) * std::max<int64_t>(1, cov_len)) == hipSuccess &&
            hipMalloc(&kb->dcounter, sizeof(uint32_t) * 64) == hipSuccess;
  for (int c = 0; c < kClasses && ok; ++c)
    ok = hipStreamCreateWithFlags(&kb->streams[c], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&kb->join[c], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&kb->fork, hipEventDisableTiming) == hipSuccess &&
       hipEventCreate(&kb->ev0) == hipSuccess && hipEventCreate(&kb->ev1) == hipSuccess;
  if (!ok) {
    kf_free(kb);
    delete kb;
    return set_error(ctx, SSM_ERR_OOM, "ssm_kf_batch_create: allocation failed (B=%d)", B);
  }
  *out = kb;
  return SSM_OK;
}

extern "C" int32_t ssm_kf_batch_run(ssm_kf_batch* kb, double* ll_out, int32_t* status_out,
                                    double* mu_out, double* Sigma_out) {
  if (!kb) return SSM_ERR_ARG;
  ssm_ctx* ctx = kb->ctx;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const bool want_state = mu_out != nullptr || Sigma_out != nullptr;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, kb->dcounter, 64);
  SSM_HIP_CHECK(ctx, hipEventRecord(kb->ev0, s));
  SSM_HIP_CHECK(ctx, hipEventRecord(kb->fork, s));
  // largest class first: it holds most of the work
  for (int c = kClasses - 1; c >= 0; --c) {
    if (kb->count[c] == 0) continue;
    hipStream_t cs = kb->streams[c];
    SSM_HIP_CHECK(ctx, hipStreamWaitEvent(cs, kb->fork, 0));
    KfBatchParams p;
    p.T = kb->T;
    p.Tf = nullptr;
    p.count = kb->count[c];
    p.order = kb->dorder + kb->start[c];
    p.counter = kb->dcounter + 16 * c;
    p.Dx = kb->dDx;
    p.Dy = kb->dDy;
    p.params = kb->dparams;
    p.param_off = kb->dpoff;
    p.ys = kb->dys;
    p.y_off = kb->dyoff;
    p.ll_out = kb->dll;
    p.status = kb->dstatus;
    p.mu_out = want_state ? kb->dmu : nullptr;
    p.Sigma_out = want_state ? kb->dSig : nullptr;
    p.mu_off = kb->dmuoff;
    p.cov_off = kb->dcovoff;
    const int rc = launch_class_nw(ctx, cs, p, c, class_nw(c));
    if (rc) return rc;
    SSM_HIP_CHECK(ctx, hipEventRecord(kb->join[c], cs));
    SSM_HIP_CHECK(ctx, hipStreamWaitEvent(s, kb->join[c], 0));
  }
  SSM_HIP_CHECK(ctx, hipEventRecord(kb->ev1, s));
  if (ll_out)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(ll_out, kb->dll, sizeof(double) * kb->B,
                                      hipMemcpyDeviceToHost, s));
  std::vector<int32_t> st(kb->B);
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(st.data(), kb->dstatus, sizeof(int32_t) * kb->B,
                                    hipMemcpyDeviceToHost, s));
  if (mu_out)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(mu_out, kb->dmu, sizeof(double) * kb->mu_len,
                                      hipMemcpyDeviceToHost, s));
  if (Sigma_out)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(Sigma_out, kb->dSig, sizeof(double) * kb->cov_len,
                                      hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  float ms = 0.f;
  SSM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, kb->ev0, kb->ev1));
  kb->last_ms = ms;
  if (status_out) std::memcpy(status_out, st.data(), sizeof(int32_t) * kb->B);
  for (int f = 0; f < kb->B; ++f)
    if (st[f] != SSM_OK)
      return set_error(ctx, SSM_ERR_NOT_PD,
                       "filter %d: innovation covariance S is not positive definite", f);
  return SSM_OK;
}

extern "C" int32_t ssm_kf_batch_tune(ssm_kf_batch* kb, int32_t T_tune, int32_t* nw_out,
                                     double* ms_out) {
  if (!kb) return SSM_ERR_ARG;
  ssm_ctx* ctx = kb->ctx;
  if (T_tune < 1) return set_error(ctx, SSM_ERR_ARG, "ssm_kf_batch_tune: T_tune must be >= 1");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int Tt = std::min(T_tune, kb->T);
  std::vector<int32_t> tt(kb->B, Tt);
  int32_t* dtf = nullptr;
  SSM_HIP_CHECK(ctx, hipMalloc(&dtf, sizeof(int32_t) * kb->B));
  SSM_HIP_CHECK(ctx, hipMemcpy(dtf, tt.data(), sizeof(int32_t) * kb->B, hipMemcpyHostToDevice));
  const int cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
  int rc = SSM_OK;
  for (int c = 0; c < kClasses && rc == SSM_OK; ++c) {
    for (int k = 0; k < 3; ++k)
      if (ms_out) ms_out[3 * c + k] = NAN;
    if (nw_out) nw_out[c] = class_nw(c);
    if (kb->count[c] == 0) continue;
    KfBatchParams p;
    p.T = Tt;
    p.Tf = dtf;
    // enough filters to fill the device several times over, costliest first
    p.count = std::min(kb->count[c], 8 * cus);
    p.order = kb->dorder + kb->start[c];
    p.counter = kb->dcounter + 16 * c;
    p.Dx = kb->dDx;
    p.Dy = kb->dDy;
    p.params = kb->dparams;
    p.param_off = kb->dpoff;
    p.ys = kb->dys;
    p.y_off = kb->dyoff;
    p.ll_out = kb->dll;
    p.status = kb->dstatus;
    p.mu_out = nullptr;
    p.Sigma_out = nullptr;
    p.mu_off = nullptr;
    p.cov_off = nullptr;
    double best = 1e300;
    int best_nw = kNwDefault[c];
    for (int k = 0; k < 3 && rc == SSM_OK; ++k) {
      const int nw = kNwCand[c][k];
      if (nw == 0) continue;
      double t_min = 1e300;
      for (int rep = 0; rep < 3 && rc == SSM_OK; ++rep) {  // first launch warms up
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, kb->dcounter, 64);
        if (hipEventRecord(kb->ev0, s) != hipSuccess) rc = SSM_ERR_HIP;
        if (rc == SSM_OK) rc = launch_class_nw(ctx, s, p, c, nw);
        if (rc == SSM_OK && (hipEventRecord(kb->ev1, s) != hipSuccess ||
                             hipEventSynchronize(kb->ev1) != hipSuccess))
          rc = SSM_ERR_HIP;
        float ms = 0.f;
        if (rc == SSM_OK && hipEventElapsedTime(&ms, kb->ev0, kb->ev1) != hipSuccess)
          rc = SSM_ERR_HIP;
        if (rc == SSM_OK && rep > 0) t_min = std::min(t_min, (double)ms);
      }
      if (ms_out) ms_out[3 * c + k] = t_min;
      if (t_min < best) {
        best = t_min;
        best_nw = nw;
      }
    }
    if (rc == SSM_OK) {
      g_nw_tuned[c].store(best_nw);
      if (nw_out) nw_out[c] = best_nw;
    }
  }
  (void)hipFree(dtf);
  if (rc == SSM_ERR_HIP) return set_error(ctx, rc, "ssm_kf_batch_tune: HIP failure");
  return rc;
}

// The filter-minor copy of a caller's array: src is B rows of P doubles (one row per
// filter), dst is P rows of B.  32 x 32 tiles through LDS (the +1 column avoids bank
// conflicts), so both the reads and the writes are coalesced runs.
__global__ void __launch_bounds__(256) k_transpose_fm(const double* __restrict__ src,
                                                      double* __restrict__ dst, int B,
                                                      int64_t P) {
  __shared__ double tile[32][33];
  const int64_t c0 = (int64_t)blockIdx.x * 32;
  const int f0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int f = f0 + j;
    const int64_t c = c0 + tx;
    if (f < B && c < P) tile[j][tx] = src[(int64_t)f * P + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int64_t c = c0 + j;
    const int f = f0 + tx;
    if (f < B && c < P) dst[c * B + f] = tile[tx][j];
  }
}

// kf_loglikelihood (GF/src/integrations/logdensity.jl:149-161) for B filters of one shape:
// the inputs are copied into one device block (grown on the context as needed) and run by
// the size class's persistent kernel in the TV form.
// gout: null, or the 8 host outputs of the gradients (mu0, Sigma0, A, b, Q, H, c, R), which
// need Dx, Dy <= 4.
static int kfll_run(ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy, int32_t T,
                    int32_t time_varying, const double* mu0, const double* Sigma0,
                    const double* A, const double* b, const double* Q, const double* H,
                    const double* c, const double* R, const double* ys, double jitter,
                    double* ll_out, int32_t* status_out, double* const* gout,
                    bool srkf = false, bool inputs_on_device = false) {
  if (!ctx) return SSM_ERR_ARG;
  if (B < 1 || T < 1 || !mu0 || !Sigma0 || !A || !b || !Q || !H || !c || !R || !ys || !ll_out)
    return set_error(ctx, SSM_ERR_ARG, "ssm_kf_loglikelihood: bad arguments");
  if (time_varying != 0 && time_varying != 1)
    return set_error(ctx, SSM_ERR_ARG, "ssm_kf_loglikelihood: time_varying must be 0 or 1");
  if (!(jitter >= 0.0))
    return set_error(ctx, SSM_ERR_ARG, "ssm_kf_loglikelihood: jitter must be >= 0");
  int cls = -1;
  for (int k = 0; k < kClasses && cls < 0; ++k)
    if (Dx >= 1 && Dy >= 1 && Dx <= kCfg[k].dx_max && Dy <= kCfg[k].dy_max) cls = k;
  if (cls < 0)
    return set_error(ctx, SSM_ERR_ARG, "ssm_kf_loglikelihood: dims Dx=%d Dy=%d outside 1..64 / 1..32",
                     Dx, Dy);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const int64_t nb = time_varying ? (int64_t)B * T : (int64_t)B;  // parameter blocks
  const int64_t nxx = (int64_t)Dx * Dx, nyx = (int64_t)Dy * Dx, nyy = (int64_t)Dy * Dy;
  const size_t n_mu0 = (size_t)B * Dx, n_S0 = (size_t)B * nxx, n_A = (size_t)(nb * nxx),
               n_b = (size_t)(nb * Dx), n_H = (size_t)(nb * nyx), n_c = (size_t)(nb * Dy),
               n_R = (size_t)(nb * nyy), n_y = (size_t)B * T * Dy;
  // with gradients: the 8 gradient arrays (the inputs' layout) and the lane kernel's
  // per-step caches; the square-root KF: the final means and factors in the first two
  const size_t n_grad = srkf ? n_mu0 + n_S0 : gout ? n_mu0 + n_S0 + 2 * n_A + n_b + n_H + n_c + n_R : 0;
  const int n_pad = std::max(Dx, Dy);
  // filters up to 4 x 4 run one per lane (inputs staged filter-minor); the reverse rule and
  // the square-root filter of larger ones run a workgroup each on the caller's layout
  const bool tiny = Dx <= kTinyD && Dy <= kTinyD;
  const bool wg = (srkf || gout) && !tiny;
  const size_t n_cache =
      (gout && !srkf && tiny) ? (size_t)T * (4 * n_pad + 4 * n_pad * n_pad) * B : 0;
  const size_t n_dbl = n_mu0 + n_S0 + 2 * n_A + n_b + n_H + n_c + n_R + n_y + (size_t)B +
                       n_grad + n_cache;
  const size_t bytes =
      (8 * n_dbl + 8 * (size_t)B /* y_off */ + 4 * 4 * (size_t)B + 256 + 255) / 256 * 256;
  // the filter-minor kernels' inputs arrive in the caller's layout behind the main block
  const size_t n_raw = n_mu0 + n_S0 + 2 * n_A + n_b + n_H + n_c + n_R + n_y;
  const size_t raw_bytes = (tiny && !inputs_on_device) ? 8 * n_raw : 0;
  int rc = ensure_device(ctx, &ctx->kfll, &ctx->kfll_bytes, bytes + raw_bytes);
  if (rc) return rc;
  // the workgroup kernels' workspace: filters in chunks of at most 2 GiB of it
  const int64_t wg_per = !wg ? 0 : srkf ? srkf_wg_doubles(Dx, Dy) : kf_grad_wg_doubles(Dx, Dy, T);
  const int wg_chunk =
      !wg ? 0 : (int)std::min<int64_t>(B, std::max<int64_t>(1, (int64_t(1) << 28) / wg_per));
  if (wg) {
    rc = ensure_device(ctx, &ctx->kfwg, &ctx->kfwg_bytes, 8 * (size_t)wg_per * wg_chunk);
    if (rc) return rc;
  }
  double* d = (double*)ctx->kfll;
  double* d_mu0 = d;
  double* d_S0 = d_mu0 + n_mu0;
  double* d_A = d_S0 + n_S0;
  double* d_b = d_A + n_A;
  double* d_Q = d_b + n_b;
  double* d_H = d_Q + n_A;
  double* d_c = d_H + n_H;
  double* d_R = d_c + n_c;
  double* d_y = d_R + n_R;
  double* d_ll = d_y + n_y;
  double* d_g = d_ll + B;           // gradients, then caches (empty without gout)
  double* d_cache = d_g + n_grad;
  int64_t* d_yoff = (int64_t*)(d_cache + n_cache);
  int32_t* d_Dx = (int32_t*)(d_yoff + B);
  int32_t* d_Dy = d_Dx + B;
  int32_t* d_order = d_Dy + B;
  int32_t* d_status = d_order + B;
  uint32_t* d_counter = (uint32_t*)(d_status + B);
  std::vector<int64_t> yoff(B);
  std::vector<int32_t> dims(3 * (size_t)B);
  for (int f = 0; f < B; ++f) {
    yoff[f] = (int64_t)f * T * Dy;
    dims[f] = Dx;
    dims[B + f] = Dy;
    dims[2 * (size_t)B + f] = f;
  }
  hipStream_t s = ctx->stream;
  // Any failure after the first asynchronous copy waits for the stream before returning, so
  // no copy is left running into or out of host memory this call owns or was handed.
  struct StreamGuard {
    hipStream_t s;
    bool armed = true;
    ~StreamGuard() {
      if (armed) (void)hipStreamSynchronize(s);
    }
  } guard{s};
  // The one-filter-per-lane kernels read their inputs filter-minor: the caller's layout is
  // uploaded unchanged behind the main block and transposed on the device.
  const bool fm = tiny;
  const int64_t nbf = time_varying ? T : 1;  // parameter blocks per filter
  double* raw = (double*)((char*)ctx->kfll + bytes);
  struct Up {
    double* dst;
    const double* src;
    int64_t per_filter;  // doubles per filter (the row length of the caller's layout)
  } ups[] = {{d_mu0, mu0, Dx},          {d_S0, Sigma0, nxx},      {d_A, A, nbf * nxx},
             {d_b, b, nbf * Dx},         {d_Q, Q, nbf * nxx},      {d_H, H, nbf * nyx},
             {d_c, c, nbf * Dy},         {d_R, R, nbf * nyy},      {d_y, ys, (int64_t)T * Dy}};
  const hipMemcpyKind in_kind = inputs_on_device ? hipMemcpyDeviceToDevice
                                                 : hipMemcpyHostToDevice;
  for (const Up& u : ups) {
    const size_t nbytes = 8 * (size_t)B * (size_t)u.per_filter;
    if (!fm) {
      SSM_HIP_CHECK(ctx, hipMemcpyAsync(u.dst, u.src, nbytes, in_kind, s));
      continue;
    }
    const double* src = u.src;  // device inputs are transposed where they lie
    if (!inputs_on_device) {
      SSM_HIP_CHECK(ctx, hipMemcpyAsync(raw, u.src, nbytes, hipMemcpyHostToDevice, s));
      src = raw;
      raw += (size_t)B * u.per_filter;
    }
    const dim3 g((unsigned)((u.per_filter + 31) / 32), (unsigned)((B + 31) / 32));
    hipLaunchKernelGGL(k_transpose_fm, g, dim3(256), 0, s, src, u.dst, B, u.per_filter);
  }
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(d_yoff, yoff.data(), 8 * (size_t)B, hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(d_Dx, dims.data(), 12 * (size_t)B, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, d_counter, 16);
  KfBatchParams p;
  p.T = T;
  p.Tf = nullptr;
  p.count = B;
  p.order = d_order;
  p.counter = d_counter;
  p.Dx = d_Dx;
  p.Dy = d_Dy;
  p.params = nullptr;
  p.param_off = nullptr;
  p.ys = d_y;
  p.y_off = d_yoff;
  p.ll_out = d_ll;
  p.status = d_status;
  p.mu_out = nullptr;
  p.Sigma_out = nullptr;
  p.mu_off = nullptr;
  p.cov_off = nullptr;
  p.jitter = jitter;
  p.tv_mu0 = d_mu0;
  p.tv_Sigma0 = d_S0;
  p.tv_A = d_A;
  p.tv_b = d_b;
  p.tv_Q = d_Q;
  p.tv_H = d_H;
  p.tv_c = d_c;
  p.tv_R = d_R;
  p.tv_blocks = time_varying ? T : 1;
  p.fm = fm ? 1 : 0;
  for (hipEvent_t& e : ctx->kfll_ev)
    if (!e) SSM_HIP_CHECK(ctx, hipEventCreate(&e));
  SSM_HIP_CHECK(ctx, hipEventRecord(ctx->kfll_ev[0], s));
  // gradient outputs in the inputs' order: mu0, Sigma0, A, b, Q, H, c, R
  const size_t g_len[8] = {n_mu0, n_S0, srkf ? 0 : n_A, srkf ? 0 : n_b, srkf ? 0 : n_A,
                           srkf ? 0 : n_H, srkf ? 0 : n_c, srkf ? 0 : n_R};
  double* g_dev[8];
  {
    double* q = d_g;
    for (int k = 0; k < 8; ++k) {
      g_dev[k] = q;
      q += gout ? g_len[k] : 0;
    }
  }
  if (wg) {
    const KfGradOut g{g_dev[0], g_dev[1], g_dev[2], g_dev[3], g_dev[4], g_dev[5], g_dev[6],
                      g_dev[7], nullptr, B};
    rc = launch_kf_wg(ctx, s, p, srkf ? nullptr : &g, g_dev[0], g_dev[1], (double*)ctx->kfwg,
                      wg_per, wg_chunk);
    if (rc) return rc;
  } else if (srkf) {
    const int n = std::max(Dx, Dy);
    auto fn = n == 1 ? k_srkf_tiny<1> : n == 2 ? k_srkf_tiny<2> : n == 3 ? k_srkf_tiny<3>
                                                                      : k_srkf_tiny<4>;
    hipLaunchKernelGGL(fn, dim3((B + kTinyThreads - 1) / kTinyThreads), dim3(kTinyThreads), 0,
                       s, p, g_dev[0], g_dev[1]);
    SSM_HIP_CHECK(ctx, hipGetLastError());
  } else if (gout) {
    KfGradOut g{g_dev[0], g_dev[1], g_dev[2], g_dev[3], g_dev[4], g_dev[5], g_dev[6],
                g_dev[7], d_cache, B};
    const int n = std::max(Dx, Dy);
    auto fn = n == 1 ? k_kf_tiny_grad<1> : n == 2 ? k_kf_tiny_grad<2>
            : n == 3 ? k_kf_tiny_grad<3> : k_kf_tiny_grad<4>;
    hipLaunchKernelGGL(fn, dim3((B + kTinyThreads - 1) / kTinyThreads), dim3(kTinyThreads), 0,
                       s, p, g);
    SSM_HIP_CHECK(ctx, hipGetLastError());
  } else if (Dx <= kTinyD && Dy <= kTinyD) {  // one filter per lane, padded to max(Dx, Dy)
    const int n = std::max(Dx, Dy);
    auto fn = n == 1 ? k_kf_tiny<1> : n == 2 ? k_kf_tiny<2> : n == 3 ? k_kf_tiny<3>
                                                                  : k_kf_tiny<kTinyD>;
    hipLaunchKernelGGL(fn, dim3((B + kTinyThreads - 1) / kTinyThreads), dim3(kTinyThreads), 0,
                       s, p);
    SSM_HIP_CHECK(ctx, hipGetLastError());
  } else {
    rc = launch_class_nw(ctx, s, p, cls, class_nw(cls));
    if (rc) return rc;
  }
  SSM_HIP_CHECK(ctx, hipEventRecord(ctx->kfll_ev[1], s));
  std::vector<int32_t> st(B);
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(ll_out, d_ll, 8 * (size_t)B, hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(st.data(), d_status, 4 * (size_t)B, hipMemcpyDeviceToHost, s));
  if (gout)
    for (int k = 0; k < 8; ++k)
      if (gout[k])
        SSM_HIP_CHECK(ctx, hipMemcpyAsync(gout[k], g_dev[k], 8 * g_len[k],
                                          hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  guard.armed = false;
  float ms = 0.f;
  SSM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->kfll_ev[0], ctx->kfll_ev[1]));
  ctx->kfll_ms = ms;
  if (status_out) std::memcpy(status_out, st.data(), 4 * (size_t)B);
  for (int f = 0; f < B; ++f)
    if (st[f] != SSM_OK)
      return set_error(ctx, SSM_ERR_NOT_PD,
                       "filter %d: innovation covariance S is not positive definite", f);
  return SSM_OK;
}

extern "C" int32_t ssm_kf_loglikelihood(ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy,
                                        int32_t T, int32_t time_varying, const double* mu0,
                                        const double* Sigma0, const double* A, const double* b,
                                        const double* Q, const double* H, const double* c,
                                        const double* R, const double* ys, double jitter,
                                        double* ll_out, int32_t* status_out) {
  return kfll_run(ctx, B, Dx, Dy, T, time_varying, mu0, Sigma0, A, b, Q, H, c, R, ys, jitter,
                  ll_out, status_out, nullptr);
}

extern "C" int32_t ssm_kf_loglikelihood_device(
    ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy, int32_t T, int32_t time_varying,
    const double* d_mu0, const double* d_Sigma0, const double* d_A, const double* d_b,
    const double* d_Q, const double* d_H, const double* d_c, const double* d_R,
    const double* d_ys, double jitter, double* ll_out, int32_t* status_out) {
  return kfll_run(ctx, B, Dx, Dy, T, time_varying, d_mu0, d_Sigma0, d_A, d_b, d_Q, d_H, d_c,
                  d_R, d_ys, jitter, ll_out, status_out, nullptr, false, true);
}

extern "C" int32_t ssm_kf_loglikelihood_grad(
    ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy, int32_t T, int32_t time_varying,
    const double* mu0, const double* Sigma0, const double* A, const double* b,
    const double* Q, const double* H, const double* c, const double* R, const double* ys,
    double jitter, double* ll_out, int32_t* status_out, double* g_mu0, double* g_Sigma0,
    double* g_A, double* g_b, double* g_Q, double* g_H, double* g_c, double* g_R) {
  if (!ctx) return SSM_ERR_ARG;
  if (Dx < 1 || Dy < 1 || Dx > kCfg[kClasses - 1].dx_max || Dy > kCfg[kClasses - 1].dy_max)
    return set_error(ctx, SSM_ERR_ARG,
                     "ssm_kf_loglikelihood_grad: gradients need 1 <= Dx <= %d, 1 <= Dy <= %d "
                     "(got %d, %d)", kCfg[kClasses - 1].dx_max, kCfg[kClasses - 1].dy_max, Dx, Dy);
  double* const gout[8] = {g_mu0, g_Sigma0, g_A, g_b, g_Q, g_H, g_c, g_R};
  return kfll_run(ctx, B, Dx, Dy, T, time_varying, mu0, Sigma0, A, b, Q, H, c, R, ys, jitter,
                  ll_out, status_out, gout);
}

extern "C" int32_t ssm_srkf_filter(ssm_ctx* ctx, int32_t B, int32_t Dx, int32_t Dy, int32_t T,
                                   int32_t time_varying, const double* mu0,
                                   const double* Sigma0, const double* A, const double* b,
                                   const double* Q, const double* H, const double* c,
                                   const double* R, const double* ys, double jitter,
                                   double* ll_out, int32_t* status_out, double* mu_out,
                                   double* U_out) {
  if (!ctx) return SSM_ERR_ARG;
  if (Dx < 1 || Dy < 1 || Dx > kCfg[kClasses - 1].dx_max || Dy > kCfg[kClasses - 1].dy_max)
    return set_error(ctx, SSM_ERR_ARG,
                     "ssm_srkf_filter: filters need 1 <= Dx <= %d, 1 <= Dy <= %d (got %d, %d)",
                     kCfg[kClasses - 1].dx_max, kCfg[kClasses - 1].dy_max, Dx, Dy);
  double* const out[8] = {mu_out, U_out, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  return kfll_run(ctx, B, Dx, Dy, T, time_varying, mu0, Sigma0, A, b, Q, H, c, R, ys, jitter,
                  ll_out, status_out, out, true);
}

extern "C" int32_t ssm_kf_loglikelihood_last_timing(ssm_ctx* ctx, double* ms) {
  if (!ctx) return SSM_ERR_ARG;
  if (ms) *ms = ctx->kfll_ms;
  return SSM_OK;
}

extern "C" int32_t ssm_kf_batch_last_timing(ssm_kf_batch* kb, double* ms) {
  if (!kb) return SSM_ERR_ARG;
  if (ms) *ms = kb->last_ms;
  return SSM_OK;
}

extern "C" int32_t ssm_kf_batch_destroy(ssm_kf_batch* kb) {
  if (!kb) return SSM_ERR_ARG;
  (void)hipSetDevice(kb->ctx->device);
  (void)hipStreamSynchronize(kb->ctx->stream);
  kf_free(kb);
  delete kb;
  return SSM_OK;
}
