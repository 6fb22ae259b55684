// mfma_team.h -- small dense linear algebra for one workgroup ("team" of NW waves) on
// matrices held in LDS: FP32 products on v_mfma_f32_16x16x4f32 and the FP64 Cholesky
// factorisation / solves of an innovation covariance by a single wave.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"

namespace ssm {

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float opget(const float* M, int ld, bool t, int r, int c) {
  return t ? M[c * ld + r] : M[r * ld + c];
}

// C (M x N) = alpha * op(A) op(B) [+ op(A2) op(B2)] [+ Add] [+ I on the first `eye`
// diagonal entries]; matrices row-major in LDS, all sizes multiples of 16.  Output 16x16
// tiles are dealt round-robin to the NW waves; each tile accumulates K/4 MFMAs, whose
// operands are fetched four products at a time.
// MFMA 16x16x4 f32 lane layout: A[l%16][l/16], B[l/16][l%16] in, C[4(l/16)+r][l%16] out.
template <int NW>
__device__ __forceinline__ void gemm16(float* C, int ldc, int M, int N, float alpha,
                                       const float* A, int lda, bool tA, const float* B,
                                       int ldb, bool tB, int K, const float* A2, int lda2,
                                       bool tA2, const float* B2, int ldb2, bool tB2, int K2,
                                       const float* Add, int ldadd, int eye) {
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) % NW;
  const int li = lane & 15, lk = lane >> 4;
  const int tn = N >> 4, ntiles = (M >> 4) * tn;
  for (int tile = wave; tile < ntiles; tile += NW) {
    const int r0 = (tile / tn) << 4, c0 = (tile % tn) << 4;
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    // K is a multiple of 16: the operands of four MFMAs are loaded before they issue, so
    // one LDS round trip covers four products (same accumulation order as one at a time)
    for (int k0 = 0; k0 < K; k0 += 16) {
      float a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = opget(A, lda, tA, r0 + li, k0 + 4 * u + lk);
        b[u] = opget(B, ldb, tB, k0 + 4 * u + lk, c0 + li);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[u], acc, 0, 0, 0);
    }
    if (alpha != 1.f) acc *= alpha;
    if (A2) {
      for (int k0 = 0; k0 < K2; k0 += 16) {
        float a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[u] = opget(A2, lda2, tA2, r0 + li, k0 + 4 * u + lk);
          b[u] = opget(B2, ldb2, tB2, k0 + 4 * u + lk, c0 + li);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[u], acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = r0 + 4 * lk + r, j = c0 + li;
      float v = acc[r];
      if (Add) v += Add[i * ldadd + j];
      if (i == j && i < eye) v += 1.f;
      C[i * ldc + j] = v;
    }
  }
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// v if every bit of m is set, +0.0 if m is zero (a branch-free, exact select).
__device__ __forceinline__ double and_mask(double v, uint32_t m) {
  return __hiloint2double(__double2hiint(v) & (int)m, __double2loint(v) & (int)m);
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = B, B+1, ..., E-1, so every
// index into a register array is a constant.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Lane S of each aligned group of W lanes (W = 1, 2 or 4), broadcast to the group (DPP
// quad_perm: the groups live inside quads; a group of one lane is the lane itself).
template <int S, int W>
__device__ __forceinline__ double group_bcast(double v) {
  static_assert(W == 1 || W == 2 || W == 4, "groups of 1, 2 or 4 lanes");
  if constexpr (W == 1) {
    return v;
  } else {
    constexpr int s0 = S, s1 = S, s2 = (W == 4) ? S : 2 + S, s3 = s2;
    constexpr int ctrl = s0 | (s1 << 2) | (s2 << 4) | (s3 << 6);
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
  }
}
// Sum over each aligned group of W lanes, in every lane of the group (fixed order).
template <int W>
__device__ __forceinline__ double group_sum(double v) {
  if constexpr (W == 1)
    return v;
  else if constexpr (W == 4)
    return (group_bcast<0, 4>(v) + group_bcast<1, 4>(v)) +
           (group_bcast<2, 4>(v) + group_bcast<3, 4>(v));
  else
    return group_bcast<0, 2>(v) + group_bcast<1, 2>(v);
}

// One wave: FP64 Cholesky S = L L' of the dy x dy matrix in Sf (FP32, row-major, ld ldSf;
// lower triangle read).  The matrix is padded to DYP with the identity, so the right-looking
// steps run in blocks of 8 without a branch inside a block (the padding factors to itself
// and leaves the leading dy x dy factor unchanged).  It lives in registers: lane l owns row
// i = l % DYP and the columns k = G m + l / DYP (G = 64 / DYP lanes per row); each step
// broadcasts the scaled pivot column through LDS and does DYP / G independent FMAs per
// lane.  Writes the padded L (lower, row-major, ld ldS) to Sd and 1 / L_jj to invd;
// colbuf holds 2 DYP doubles.  Returns 0 if S is not positive definite.
// *logdet = sum_j 2 log L_jj, in a fixed order.
template <int DYP>
__device__ __forceinline__ int chol_factor_wave(const float* Sf, int ldSf, int dy, double* Sd, int ldS,
                                double* colbuf, double* invd, double* logdet) {
  static_assert(DYP == 8 || DYP == 16 || DYP == 32, "DYP must be 8, 16 or 32");
  constexpr int G = 64 / DYP, E = DYP / G;
  const int lane = threadIdx.x & 63;
  const int i = lane % DYP, h = lane / DYP;
  double a[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = G * m + h;
    a[m] = (i < dy && k <= i) ? (double)Sf[i * ldSf + k] : ((i == k) ? 1.0 : 0.0);
  }
  bool good = true;
  static_for<0, DYP / 8>([&](auto J0) {
    constexpr int j0 = 8 * decltype(J0)::value;
    if (j0 < dy) {  // whole blocks of 8 padding steps are skipped (uniform branch)
      static_for<0, 8>([&](auto JJ) {
        constexpr int j = j0 + decltype(JJ)::value;
        constexpr int mj = j / G, hj = j % G;
        const double ajj = readlane_d(a[mj], hj * DYP + j);
        good = good && (ajj > 0.0);
        const double d = sqrt(ajj > 0.0 ? ajj : 1.0);
        const double inv = 1.0 / d;
        // the lanes holding column j scale it; the others write their copy to the spare
        // half of colbuf, so the store needs no branch
        const double v = (i > j) ? a[mj] * inv : ((i == j) ? d : 0.0);
        const bool own = (h == hj);
        a[mj] = own ? v : a[mj];
        colbuf[(own ? 0 : DYP) + i] = (i > j) ? v : 0.0;
        invd[j] = inv;
        wave_lds_sync();
        const double li = colbuf[i];
        // trailing update of the columns k = G m + h > j.  Entries above the diagonal
        // (k > i) are updated too but never read, so only k > j needs a mask: entries with
        // G m + G - 1 <= j are skipped at compile time, G m > j needs no mask, and the one
        // straddling entry keeps the lanes with h > j - G m.
        static_for<0, E>([&](auto M) {
          constexpr int m = decltype(M)::value, dd = j - G * m;
          if constexpr (dd < G - 1) {
            double lk = colbuf[G * m + h];
            if constexpr (dd >= 0) lk = and_mask(lk, h > dd ? 0xffffffffu : 0u);
            a[m] -= li * lk;
          }
        });
        wave_lds_sync();
      });
    }
  });
#pragma unroll
  for (int m = 0; m < E; ++m) {  // the whole padded square: the upper part is zero
    const int k = G * m + h;
    Sd[i * ldS + k] = (k <= i) ? a[m] : 0.0;
  }
  wave_lds_sync();
  double lg = (lane < dy && good) ? 2.0 * log(Sd[lane * ldS + lane]) : 0.0;
  lg = wave_sum_dpp(lg);
  *logdet = lg;
  return good ? 1 : 0;
}

// Triangular solves with the padded factor of chol_factor_wave, one right-hand side per
// group of DYP / 8 lanes (lane r of a group holds rows 8r .. 8r+7, so 16 FP64 registers
// per lane): column `col` < dx of P (FP32, ld ldP) gives row-major KT[:, col] =
// S^-1 P[:, col] (zero-padded to dyp rows); col == dx takes z and returns z' S^-1 z to
// every lane of its group; columns dx..dxp-1 of KT are zero padding.  Axpy order: at step
// k the owner of row k scales it by 1 / L_kk, a DPP broadcast hands it to the group, and
// each lane updates its rows.  Row blocks of 8 past dy are skipped (exact zeros); r is
// this lane's index in its group.  All lanes of a group must call together.
template <int DYP>
__device__ __forceinline__ double solve_group(int col, int r, const double* Sd, int ldS,
                                              const double* invd, const float* P, int ldP,
                                              const float* zv, float* KT, int ldK, int dx,
                                              int dxp, int dy, int dyp) {
  static_assert(DYP == 8 || DYP == 16 || DYP == 32, "DYP must be 8, 16 or 32");
  if (col > dx) {
    if (col < dxp)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (8 * r + j < dyp) KT[(8 * r + j) * ldK + col] = 0.f;
    return 0.0;
  }
  double x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int q = 8 * r + j;  // q < DYP: P and zv hold DYP rows, so the loads stay in bounds
    const double pv = (double)P[q * ldP + (col < dx ? col : 0)];
    const double zq = (double)zv[q];
    x[j] = (q < dy) ? ((col < dx) ? pv : zq) : 0.0;
  }
  // forward: L y = b (compile-time row ranges, bit masks and register selects instead of
  // branches; every load is in bounds and finite, the factor's whole padded square being
  // written)
  static_for<0, DYP / 8>([&](auto K0) {
    constexpr int rb = decltype(K0)::value;  // row block owning the pivots
    if (8 * rb < dy) {
      static_for<0, 8>([&](auto KK) {
        constexpr int kk = decltype(KK)::value, k = 8 * rb + kk;
        // every lane scales its slot kk; the owner's value is broadcast and kept, the
        // others keep theirs (a register select, no branch)
        const double xk = group_bcast<rb, DYP / 8>(x[kk] * invd[k]);
        x[kk] = (r == rb) ? xk : x[kk];
        // rows 8 r + j > k: r > c with c = floor((k - j) / 8) known at compile time
        static_for<0, 8>([&](auto J) {
          constexpr int j = decltype(J)::value;
          constexpr int c = (k - j >= 0) ? (k - j) / 8 : -1;
          if constexpr (c < DYP / 8 - 1) {
            double l = Sd[(8 * r + j) * ldS + k];
            if constexpr (c >= 0) l = and_mask(l, r > c ? 0xffffffffu : 0u);
            x[j] -= l * xk;
          }
        });
      });
    }
  });
  if (col == dx) {
    if (dx < dxp)  // column dx is also padding of KT
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (8 * r + j < dyp) KT[(8 * r + j) * ldK + col] = 0.f;
    double qd = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) qd += x[j] * x[j];
    return group_sum<DYP / 8>(qd);
  }
  // backward: L' x = y
  static_for<0, DYP / 8>([&](auto K0) {
    constexpr int rb = DYP / 8 - 1 - decltype(K0)::value;
    if (8 * rb < dy) {
      static_for<0, 8>([&](auto KK) {
        constexpr int kk = 7 - decltype(KK)::value, k = 8 * rb + kk;
        const double xk = group_bcast<rb, DYP / 8>(x[kk] * invd[k]);
        x[kk] = (r == rb) ? xk : x[kk];
        // rows 8 r + j < k: r <= c with c = floor((k - j - 1) / 8) known at compile time
        static_for<0, 8>([&](auto J) {
          constexpr int j = decltype(J)::value;
          constexpr int c = (k - j - 1 >= 0) ? (k - j - 1) / 8 : -1;
          if constexpr (c >= 0) {
            double l = Sd[k * ldS + 8 * r + j];
            if constexpr (c < DYP / 8 - 1) l = and_mask(l, r <= c ? 0xffffffffu : 0u);
            x[j] -= l * xk;
          }
        });
      });
    }
  });
#pragma unroll
  for (int j = 0; j < 8; ++j) {  // KT holds DYP rows; rows past dy get zeros
    const int q = 8 * r + j;
    KT[q * ldK + col] = (q < dy) ? (float)x[j] : 0.f;
  }
  return 0.0;
}

// One wave: the FP64 Cholesky of the dy x dy innovation covariance in Sf (FP32, ld ldSf)
// into Sd (ld ldS), then KT (dyp x dxp, FP32) = S^-1 P for the dx columns of P (FP32, ld
// ldP), zero-padded, *quad = z' S^-1 z and *logdet = log det S.  work holds 3 DYP doubles.
// Returns 0 if S is not positive definite.  DYP bounds dy at compile time.
template <int DYP>
__device__ int chol_solve_wave(double* Sd, int ldS, const float* Sf, int ldSf, const float* P,
                               int ldP, const float* zv, float* KT, int ldK, int dx, int dxp,
                               int dy, int dyp, double* quad, double* logdet, double* work) {
  const int lane = threadIdx.x & 63;
  double* colbuf = work;  // 2 DYP
  double* invd = work + 2 * DYP;
  const int good = chol_factor_wave<DYP>(Sf, ldSf, dy, Sd, ldS, colbuf, invd, logdet);
  constexpr int LPC = DYP / 8;  // lanes per right-hand side
  double qd = 0.0;
  if (good) {
    const int ncol = dx + 1 > dxp ? dx + 1 : dxp;
    const int r = lane % LPC;
    for (int c0 = 0; c0 < ncol; c0 += 64 / LPC) {  // all lanes take part in every pass
      const int col = c0 + lane / LPC;
      const double q = solve_group<DYP>(col < ncol ? col : dxp, r, Sd, ldS, invd, P, ldP, zv,
                                        KT, ldK, dx, dxp, dy, dyp);
      if (col == dx) qd = q;
    }
  }
  // the group that solved column dx holds the quadratic form; broadcast it
  qd = __shfl(qd, ((dx * LPC) & 63), 64);
  *quad = qd;
  wave_lds_sync();
  return good;
}

}  // namespace ssm
