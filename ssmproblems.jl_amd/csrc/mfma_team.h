// mfma_team.h -- small dense linear algebra for one workgroup ("team" of NW waves) on
// matrices held in LDS: FP32 products on v_mfma_f32_16x16x4f32 and the FP64 Cholesky
// factorisation / solves of an innovation covariance by a single wave.
#pragma once
#include <hip/hip_runtime.h>

namespace ssm {

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float opget(const float* M, int ld, bool t, int r, int c) {
  return t ? M[c * ld + r] : M[r * ld + c];
}

// C (M x N) = alpha * op(A) op(B) [+ op(A2) op(B2)] [+ Add] [+ I on the first `eye`
// diagonal entries]; matrices row-major in LDS, all sizes multiples of 16.  Output 16x16
// tiles are dealt round-robin to the NW waves; each tile accumulates K/4 MFMAs.
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
    for (int k0 = 0; k0 < K; k0 += 4) {
      const float a = opget(A, lda, tA, r0 + li, k0 + lk);
      const float b = opget(B, ldb, tB, k0 + lk, c0 + li);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
    if (alpha != 1.f) acc *= alpha;
    if (A2) {
      for (int k0 = 0; k0 < K2; k0 += 4) {
        const float a = opget(A2, lda2, tA2, r0 + li, k0 + lk);
        const float b = opget(B2, ldb2, tB2, k0 + lk, c0 + li);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
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

// Orders this wave's LDS traffic: DS instructions of one wave complete in issue order, so
// a compiler-level fence at wavefront scope is enough between dependent phases.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One wave: FP64 Cholesky S = L L' in place (lower triangle of Sd, row-major, ld = ldS) of
// the dy x dy innovation covariance read from Sf (FP32, ld = ldSf); then
//   KT (dyp x dxp, FP32) = S^-1 P  for the dx columns of P (FP32, ld = ldP), zero-padded,
//   *quad = z' S^-1 z (FP64),  *logdet = log det S.
// Returns 0 if S is not positive definite.  DYP bounds dy at compile time.
template <int DYP>
__device__ int chol_solve_wave(double* Sd, int ldS, const float* Sf, int ldSf, const float* P,
                               int ldP, const float* zv, float* KT, int ldK, int dx, int dxp,
                               int dy, int dyp, double* quad, double* logdet) {
  const int lane = threadIdx.x & 63;
  for (int e = lane; e < dy * dy; e += 64) {
    const int i = e / dy, j = e % dy;
    Sd[i * ldS + j] = (double)Sf[i * ldSf + j];
  }
  wave_lds_sync();
  int good = 1;
  double ld = 0.0;
  for (int j = 0; j < dy; ++j) {
    double d = Sd[j * ldS + j];
    if (!(d > 0.0)) {
      good = 0;
      break;
    }
    d = sqrt(d);
    ld += 2.0 * log(d);
    wave_lds_sync();
    if (lane == 0) Sd[j * ldS + j] = d;
    for (int i = j + 1 + lane; i < dy; i += 64) Sd[i * ldS + j] /= d;
    wave_lds_sync();
    for (int i = j + 1 + lane; i < dy; i += 64) {
      const double lij = Sd[i * ldS + j];
      for (int k = j + 1; k <= i; ++k) Sd[i * ldS + k] -= lij * Sd[k * ldS + j];
    }
    wave_lds_sync();
  }
  double qd = 0.0;
  if (good) {
    // columns 0..dx-1 of P give K' = S^-1 P; column dx is z (forward solve only).
    for (int col = lane; col <= dx; col += 64) {
      double x[DYP];
#pragma unroll
      for (int q = 0; q < DYP; ++q)
        x[q] = (q < dy) ? (double)((col < dx) ? P[q * ldP + col] : zv[q]) : 0.0;
#pragma unroll
      for (int q = 0; q < DYP; ++q) {
        if (q < dy) {
          double s = x[q];
#pragma unroll
          for (int k = 0; k < q; ++k) s -= Sd[q * ldS + k] * x[k];
          x[q] = s / Sd[q * ldS + q];
        }
      }
      if (col == dx) {
#pragma unroll
        for (int q = 0; q < DYP; ++q)
          if (q < dy) qd += x[q] * x[q];
      } else {
#pragma unroll
        for (int q = DYP - 1; q >= 0; --q) {
          if (q < dy) {
            double s = x[q];
#pragma unroll
            for (int k = q + 1; k < DYP; ++k)
              if (k < dy) s -= Sd[k * ldS + q] * x[k];
            x[q] = s / Sd[q * ldS + q];
          }
        }
#pragma unroll
        for (int q = 0; q < DYP; ++q)
          if (q < dyp) KT[q * ldK + col] = (q < dy) ? (float)x[q] : 0.f;
      }
    }
    for (int e = lane; e < dyp * (dxp - dx); e += 64) {
      const int q = e / (dxp - dx), col = dx + e % (dxp - dx);
      KT[q * ldK + col] = 0.f;
    }
  }
  // the lane that solved column dx holds the quadratic form; broadcast it
  const int zlane = dx & 63;
  qd = __shfl(qd, zlane, 64);
  *quad = qd;
  *logdet = ld;
  wave_lds_sync();
  return good;
}

}  // namespace ssm
