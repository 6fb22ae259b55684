// kf_team.h -- dense linear algebra for the Kalman kernels (kf.hip, rbpf.hip) on gfx950:
// swizzled LDS matrices read as 16-byte MFMA operands, FP32 matrix-core tiles, and the FP64
// inverse of an innovation covariance by one wave (Gauss-Jordan blocks exchanged through
// DPP and permlanes, the Schur complement on the FP64 matrix cores).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include <type_traits>

#include "device_common.h"

namespace ssm {

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double readlane_d(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
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

// LDS matrices are row-major with LD columns and no padding; the 16-byte column group
// c / 4 of row r is XOR-ed with r mod min(LD / 4, 16).  A lane reading four consecutive
// columns of its row (one ds_read_b128) and the transposed write of an MFMA output tile
// then spread a wave's rows over distinct banks.
template <int LD>
__device__ __forceinline__ int swz(int r, int c) {
  constexpr int M = (LD / 4 > 16 ? 16 : LD / 4) - 1;
  return r * LD + (c ^ ((r & M) << 2));
}

__device__ __forceinline__ f4v mfma4(f4v a, f4v b, f4v acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], acc, 0, 0, 0);
  return acc;
}

// The 16 x 16 tile at (r0, c0) of X Y over k < K, X row-major (LDX columns) and Y given as
// Y^T row-major (LDY columns).  MFMA u of the k-block k0 takes k = k0 + 4 lk + u in lane
// (li, lk), so each operand is one 16-byte read per lane and block.
// MFMA 16x16x4 f32 lane layout: A[l%16][l/16], B[l/16][l%16] in, C[4(l/16)+r][l%16] out.
template <int LDX, int LDY>
__device__ __forceinline__ f4v kf_mm(const float* X, const float* YT, int r0, int c0, int K, int li, int lk) {
  constexpr int KB = (LDX < LDY ? LDX : LDY) / 16;
  f4v a[KB], b[KB];  // every block's operands are loaded before the first product
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    if (16 * kb < K) {
      a[kb] = *reinterpret_cast<const f4v*>(X + swz<LDX>(r0 + li, 16 * kb + 4 * lk));
      b[kb] = *reinterpret_cast<const f4v*>(YT + swz<LDY>(c0 + li, 16 * kb + 4 * lk));
    }
  }
  f4v acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < KB; ++kb)
    if (16 * kb < K) acc = mfma4(a[kb], b[kb], acc);
  return acc;
}

// Two tiles of X Y at (r0, c0) and (r1, c1), their matrix-core chains interleaved so each
// hides the other's dependent-issue latency.
template <int LDX, int LDY>
__device__ __forceinline__ void kf_mm2(const float* X, const float* YT, int r0, int c0, int r1,
                                       int c1, int K, int li, int lk, f4v& acc0, f4v& acc1) {
  constexpr int KB = (LDX < LDY ? LDX : LDY) / 16;
  // operands of block kb + 1 are loaded while block kb's products issue
  f4v a0 = *reinterpret_cast<const f4v*>(X + swz<LDX>(r0 + li, 4 * lk));
  f4v b0 = *reinterpret_cast<const f4v*>(YT + swz<LDY>(c0 + li, 4 * lk));
  f4v a1 = *reinterpret_cast<const f4v*>(X + swz<LDX>(r1 + li, 4 * lk));
  f4v b1 = *reinterpret_cast<const f4v*>(YT + swz<LDY>(c1 + li, 4 * lk));
  acc0 = f4v{0.f, 0.f, 0.f, 0.f};
  acc1 = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    if (16 * kb < K) {
      f4v na0 = a0, nb0 = b0, na1 = a1, nb1 = b1;
      if (kb + 1 < KB && 16 * (kb + 1) < K) {
        na0 = *reinterpret_cast<const f4v*>(X + swz<LDX>(r0 + li, 16 * (kb + 1) + 4 * lk));
        nb0 = *reinterpret_cast<const f4v*>(YT + swz<LDY>(c0 + li, 16 * (kb + 1) + 4 * lk));
        na1 = *reinterpret_cast<const f4v*>(X + swz<LDX>(r1 + li, 16 * (kb + 1) + 4 * lk));
        nb1 = *reinterpret_cast<const f4v*>(YT + swz<LDY>(c1 + li, 16 * (kb + 1) + 4 * lk));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[u], b0[u], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[u], b1[u], acc1, 0, 0, 0);
      }
      a0 = na0;
      b0 = nb0;
      a1 = na1;
      b1 = nb1;
    }
  }
}

// The tiles t = first, first + stride, ... < count of X Y, two at a time; tile(t, r0, c0)
// names tile t's origin and out(r0, c0, acc) consumes its product.
template <int LDX, int LDY, typename TileF, typename OutF>
__device__ __forceinline__ void kf_mm_tiles(const float* X, const float* YT, int K, int first,
                                            int stride, int count, int li, int lk, TileF tile,
                                            OutF out) {
  for (int t = first; t < count; t += 2 * stride) {
    int r0, c0;
    tile(t, r0, c0);
    if (t + stride < count) {
      int r1, c1;
      tile(t + stride, r1, c1);
      f4v acc0, acc1;
      kf_mm2<LDX, LDY>(X, YT, r0, c0, r1, c1, K, li, lk, acc0, acc1);
      out(r0, c0, acc0);
      out(r1, c1, acc1);
    } else {
      out(r0, c0, kf_mm<LDX, LDY>(X, YT, r0, c0, K, li, lk));
    }
  }
}

// Stores an output tile at (r0, c0): as is (lane writes its 4 rows of one column), or
// transposed into C^T (one 16-byte write per lane).
template <int LD>
__device__ __forceinline__ void kf_store(float* C, int r0, int c0, f4v v, int li, int lk) {
#pragma unroll
  for (int e = 0; e < 4; ++e) C[swz<LD>(r0 + 4 * lk + e, c0 + li)] = v[e];
}
template <int LD>
__device__ __forceinline__ void kf_store_t(float* CT, int r0, int c0, f4v v, int li, int lk) {
  *reinterpret_cast<f4v*>(CT + swz<LD>(c0 + li, r0 + 4 * lk)) = v;
}
// The lane's 4 entries of the tile at (r0, c0) of a matrix stored transposed.
template <int LD>
__device__ __forceinline__ f4v kf_load_t(const float* CT, int r0, int c0, int li, int lk) {
  return *reinterpret_cast<const f4v*>(CT + swz<LD>(c0 + li, r0 + 4 * lk));
}

// ---- FP64 inverse of the innovation covariance (one wave) -----------------------------
// 16 x 16 blocks are held "column-strided": lane l has row i = l % 16, columns h + 4 e
// (h = l / 16, e = 0..3), which is the output layout of v_mfma_f64_16x16x4f64.

// Lane S of each row of 16 lanes, broadcast to that row (DPP row_newbcast).
template <int S>
__device__ __forceinline__ double row_bcast16(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x150 + S, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x150 + S, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
// Row R of 16 lanes broadcast to all four rows, lane for lane: permlane16_swap of v with
// itself gives rows (0, 0, 2, 2) and (1, 1, 3, 3), and permlane32_swap of the one holding
// row R with itself gives (R, R, R, R).
template <int R>
__device__ __forceinline__ uint32_t rows_bcast(uint32_t v) {
  const auto s16 = __builtin_amdgcn_permlane16_swap(v, v, false, false);
  const uint32_t w = (R & 1) ? s16[1] : s16[0];
  const auto s32 = __builtin_amdgcn_permlane32_swap(w, w, false, false);
  return (R & 2) ? s32[1] : s32[0];
}
template <int R>
__device__ __forceinline__ double rows_bcast_d(double v) {
  return __hiloint2double((int)rows_bcast<R>((uint32_t)__double2hiint(v)),
                          (int)rows_bcast<R>((uint32_t)__double2loint(v)));
}

// One wave: in-place Gauss-Jordan inverse in FP64 of the leading n x n block of a 16 x 16
// symmetric positive definite matrix M (column-strided), without pivoting.  Step k: the
// pivot p = a_kk (readlane), row k to each row of 16 lanes (DPP), column k to the four rows
// (permlane swaps), then a_ij -= a_ik a_kj / p, a_kj = a_kj / p, a_ik = -a_ik / p,
// a_kk = 1 / p (row k's lanes start from zero, so one update form serves every row).  Rows
// and columns >= n are zero and stay zero.  The pivots are those of the LDL' factorisation,
// all positive iff M is positive definite; lane k keeps pivot k in *piv.  Returns false if
// one is not.
__device__ __forceinline__ bool gj16(double a[4], int n, double* piv) {
  const int lane0 = threadIdx.x & 63;
  bool good = true;
  static_for<0, 16>([&](auto KK) {
    constexpr int k = decltype(KK)::value, ek = k >> 2, hk = k & 3;
    if (k < n) {  // uniform
      // the lane index passes through empty asm each step, so the step's lane masks are
      // formed where they are used rather than all sixteen steps' kept live (and spilled)
      int lane = lane0;
      asm volatile("" : "+v"(lane));
      const int i = lane & 15, h = lane >> 4;
      const double pk = readlane_d(a[ek], k + 16 * hk);
      double rk[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) rk[e] = row_bcast16<k>(a[e]);
      const double cik = rows_bcast_d<hk>(a[ek]);
      good = good && (pk > 0.0);
      const double pp = pk > 0.0 ? pk : 1.0;
      // 1 / pp: the hardware reciprocal and two Newton steps (error below 1 ulp)
      double r = __builtin_amdgcn_rcp(pp);
      r = fma(r, fma(-pp, r, 1.0), r);
      r = fma(r, fma(-pp, r, 1.0), r);
      *piv = (lane == k) ? pk : *piv;
      const bool rowk = (i == k);
      const double c = rowk ? -r : cik * r;
#pragma unroll
      for (int e = 0; e < 4; ++e) a[e] = fma(-c, rk[e], rowk ? 0.0 : a[e]);
      const double colk = rowk ? r : -cik * r;
      a[ek] = (h == hk) ? colk : a[ek];
    }
  });
  return good;
}

// Q P' for 16 x 16 FP64 matrices, all column-strided, on the matrix cores: MFMA u takes
// k = h + 4 u in lane (i, h), so P's and Q's own registers are its two operands.
__device__ __forceinline__ void mm16_qpt(const double p[4], const double q[4], double out[4]) {
  typedef double d4v __attribute__((ext_vector_type(4)));
  d4v acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(p[u], q[u], acc, 0, 0, 0);
#pragma unroll
  for (int e = 0; e < 4; ++e) out[e] = acc[e];
}

// acc + Q P' (mm16_qpt accumulating).
__device__ __forceinline__ void mm16_qpt_acc(const double p[4], const double q[4], double acc[4]) {
  typedef double d4v __attribute__((ext_vector_type(4)));
  d4v a = {acc[0], acc[1], acc[2], acc[3]};
#pragma unroll
  for (int u = 0; u < 4; ++u) a = __builtin_amdgcn_mfma_f64_16x16x4f64(p[u], q[u], a, 0, 0, 0);
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = a[e];
}

// S^-1 in FP64 registers, as column-strided 16 x 16 blocks: s11 = (S^-1)_11, s12 = (S^-1)_12,
// s21 = (S^-1)_21 = s12' and s22 = (S^-1)_22 (the last three zero when dy <= 16).  The
// layout is the same in every wave, so a copy through LDS indexed by lane hands it over.
struct SpdInv {
  double s11[4], s12[4], s21[4], s22[4];
};

// One wave: the FP64 inverse of the dy x dy symmetric positive definite matrix S (FP32,
// swizzled with LD = DYP, exactly symmetric).  Up to 16: gj16.  Beyond, by blocks
// S = [A B; B' D]: X = A^-1 B, C = D - B' X (the Schur complement), Y = X C^-1,
// S^-1 = [A^-1 + Y X', -Y; -Y', C^-1], log det S = log det A + log det C, the products on
// the FP64 matrix cores with the blocks' own registers as operands.  *quad = z' S^-1 z
// and *logdet from the pivots (fixed-order DPP sums).  Returns 0 if S is not positive
// definite.
template <int DYP>
__device__ __forceinline__ int spd_inverse_wave(const float* S, int dy, const float* zv,
                                                SpdInv& v, double* logdet, double* quad) {
  static_assert(DYP == 16 || DYP == 32, "DYP must be 16 or 32");
  const int lane = threadIdx.x & 63, i = lane & 15, h = lane >> 4;
  auto sget = [&](int r, int c) -> double {
    return (r < dy && c < dy) ? (double)S[swz<DYP>(r, c)] : 0.0;
  };
  auto zget = [&](int r) -> double { return r < dy ? (double)zv[r] : 0.0; };
  double pa = 1.0, pc = 1.0;
#pragma unroll
  for (int e = 0; e < 4; ++e) v.s11[e] = sget(i, h + 4 * e);
  bool good = gj16(v.s11, dy < 16 ? dy : 16, &pa);
  double q = 0.0;
  if (DYP == 32 && dy > 16) {  // uniform
    double bt[4], x[4], xt[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) bt[e] = sget(16 + i, h + 4 * e);  // B' (S symmetric)
    mm16_qpt(bt, v.s11, x);    // X = A^-1 B
    mm16_qpt(v.s11, bt, xt);   // X' = B' A^-1
    mm16_qpt(xt, bt, v.s22);   // B' X
#pragma unroll
    for (int e = 0; e < 4; ++e) v.s22[e] = sget(16 + i, 16 + h + 4 * e) - v.s22[e];
    good = gj16(v.s22, dy - 16, &pc) && good;  // C^-1
    mm16_qpt(v.s22, x, v.s12);  // Y = X C^-1
    mm16_qpt(x, v.s22, v.s21);  // Y' = C^-1 X'
    double yx[4];
    mm16_qpt(x, v.s12, yx);     // Y X'
    const double z1i = zget(i), z2i = zget(16 + i);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = h + 4 * e;
      v.s11[e] += yx[e];
      v.s12[e] = -v.s12[e];
      v.s21[e] = -v.s21[e];
      q += z1i * v.s11[e] * zget(c) + 2.0 * z1i * v.s12[e] * zget(16 + c) +
           z2i * v.s22[e] * zget(16 + c);
    }
  } else {
    const double zi = zget(i);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      q += zi * v.s11[e] * zget(h + 4 * e);
      v.s12[e] = v.s21[e] = v.s22[e] = 0.0;
    }
  }
  *quad = wave_sum_dpp(q);
  // lane k < 16 holds pivot k of each block (1 past the block's size)
  *logdet = wave_sum_dpp(good && lane < 16 ? log(pa * pc) : 0.0);
  return good ? 1 : 0;
}

// One wave: row block r of the gain K = P' S^-1 (16 rows of dy) in FP64 on the matrix
// cores, from P' (FP32, swizzled, LD = DYP) into K (FP32, same layout):
// [P'_r1 P'_r2] S^-1 = [P'_r1 S11 + P'_r2 S21, P'_r1 S12 + P'_r2 S22].  The inverse never
// leaves FP64, so K carries only its own rounding however S is conditioned.
template <int DYP>
__device__ __forceinline__ void gain_rows(const SpdInv& v, int dy, const float* PT, float* K,
                                          int r) {
  const int lane = threadIdx.x & 63, i = lane & 15, h = lane >> 4;
  double p1[4], k1[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int e = 0; e < 4; ++e) p1[e] = (double)PT[swz<DYP>(16 * r + i, h + 4 * e)];
  mm16_qpt_acc(v.s11, p1, k1);
  if (DYP == 32 && dy > 16) {  // uniform
    double p2[4], k2[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int e = 0; e < 4; ++e) p2[e] = (double)PT[swz<DYP>(16 * r + i, 16 + h + 4 * e)];
    mm16_qpt_acc(v.s12, p2, k1);
    mm16_qpt_acc(v.s21, p1, k2);
    mm16_qpt_acc(v.s22, p2, k2);
#pragma unroll
    for (int e = 0; e < 4; ++e) K[swz<DYP>(16 * r + i, 16 + h + 4 * e)] = (float)k2[e];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) K[swz<DYP>(16 * r + i, h + 4 * e)] = (float)k1[e];
}

// The inverse's registers through LDS (NB blocks of 4 doubles a lane, lane-contiguous).
template <int NB>
__device__ __forceinline__ void spd_inv_store(const SpdInv& v, double* buf) {
  const int lane = threadIdx.x & 63;
  const double* src[4] = {v.s11, v.s12, v.s21, v.s22};
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int e = 0; e < 4; ++e) buf[(4 * b + e) * 64 + lane] = src[b][e];
}
template <int NB>
__device__ __forceinline__ void spd_inv_load(SpdInv& v, const double* buf) {
  const int lane = threadIdx.x & 63;
  double* dst[4] = {v.s11, v.s12, v.s21, v.s22};
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int e = 0; e < 4; ++e) dst[b][e] = b < NB ? buf[(4 * b + e) * 64 + lane] : 0.0;
}

}  // namespace ssm
