// device_common.h -- device helpers shared by the gfx950 kernels of the filtering library:
// counter-based Philox4x32-10 noise, uniform conversions, Box-Muller, and wave64 / block
// reductions and scans.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssm {

constexpr int kWave = 64;

// Philox stream ids (word 2 of the counter): independent noise streams per purpose.
constexpr uint32_t kStreamPropagate = 0u;
constexpr uint32_t kStreamResample = 1u;
constexpr uint32_t kStreamInit = 2u;
constexpr uint32_t kStreamOuter = 3u;
constexpr uint32_t kStreamAncestor = 4u;  // conditional SMC: the reference's ancestor draw

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter {c0..c3}, key {k0, k1}.
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                            uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // one 32x32->64 product per multiplier (a single v_mad_u64_u32 each)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return U4{c0, c1, c2, c3};
}

// 24-bit uniforms as float: (0,1] and [0,1).
__device__ __forceinline__ float u24_open0(uint32_t w) {
  return (float)((w >> 8) + 1u) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ float u24_closed0(uint32_t w) {
  return (float)(w >> 8) * (1.0f / 16777216.0f);
}
// 53-bit uniforms as double from two words: (0,1] and [0,1).
__device__ __forceinline__ double u53_closed0(uint32_t a, uint32_t b) {
  const uint64_t v = (((uint64_t)a << 32) | b) >> 11;
  return (double)v * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double u53_open0(uint32_t a, uint32_t b) {
  const uint64_t v = (((uint64_t)a << 32) | b) >> 11;
  return (double)(v + 1u) * (1.0 / 9007199254740992.0);
}

// Box-Muller in FP32 on the hardware transcendental unit: v_log_f32 is log2, and
// v_sin_f32 / v_cos_f32 take their argument in revolutions, so sin(2 pi u) is one op.
__device__ __forceinline__ void box_muller_f32(uint32_t w0, uint32_t w1, float& z0,
                                               float& z1) {
  const float u1 = u24_open0(w0), u2 = u24_closed0(w1);
  // -2 ln u1 = -2 ln2 * log2(u1)
  const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
  z0 = r * __builtin_amdgcn_cosf(u2);
  z1 = r * __builtin_amdgcn_sinf(u2);
}

__device__ __forceinline__ void box_muller_f64(uint32_t a, uint32_t b, uint32_t c,
                                               uint32_t d, double& z0, double& z1) {
  const double u1 = u53_open0(a, b), u2 = u53_closed0(c, d);
  const double r = sqrt(-2.0 * log(u1));
  double s, co;
  sincospi(2.0 * u2, &s, &co);
  z0 = r * co;
  z1 = r * s;
}

// Standard normals for component block j of particle i (f32 layout: 4 per Philox call).
__device__ __forceinline__ void normals4_f32(uint64_t seed, uint32_t stream, uint32_t t,
                                             uint32_t i, uint32_t j, float z[4]) {
  const U4 w = philox4x32_10(i, t, stream, j, (uint32_t)seed, (uint32_t)(seed >> 32));
  box_muller_f32(w.x, w.y, z[0], z[1]);
  box_muller_f32(w.z, w.w, z[2], z[3]);
}
// f64 layout: 2 per Philox call.
__device__ __forceinline__ void normals2_f64(uint64_t seed, uint32_t stream, uint32_t t,
                                             uint32_t i, uint32_t j, double z[2]) {
  const U4 w = philox4x32_10(i, t, stream, j, (uint32_t)seed, (uint32_t)(seed >> 32));
  box_muller_f64(w.x, w.y, w.z, w.w, z[0], z[1]);
}
// Resampling uniform u_k in [0,1) for step t.
__device__ __forceinline__ double resample_uniform(uint64_t seed, uint32_t t, uint32_t k) {
  const U4 w = philox4x32_10(k, t, kStreamResample, 0u, (uint32_t)seed,
                             (uint32_t)(seed >> 32));
  return u53_closed0(w.x, w.y);
}

// A [0, 1) uniform from any stream, keyed like resample_uniform.
__device__ __forceinline__ double stream_uniform(uint64_t seed, uint32_t stream, uint32_t t,
                                                uint32_t k) {
  const U4 w = philox4x32_10(k, t, stream, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
  return u53_closed0(w.x, w.y);
}

// ---- wave64 primitives ------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_shfl_xor(T v, int m) {
  return __shfl_xor(v, m, kWave);
}
template <typename T>
__device__ __forceinline__ T wave_shfl_up(T v, int d) {
  return __shfl_up(v, d, kWave);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += wave_shfl_xor(v, m);
  return v;
}

// Inclusive prefix sum across the 64 lanes (Hillis-Steele, 6 shuffle steps).
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const T o = wave_shfl_up(v, d);
    if (lane >= d) v += o;
  }
  return v;
}

// A store at device scope: it writes through the XCD's L2 instead of leaving a dirty line
// there.  For a kernel's bulk outputs, which the next kernel reads from every XCD, this
// lets the writes drain while the kernel runs rather than in the write-back at its end
// (measured 0.7 us less per bootstrap-filter step at N = 2^20).
template <typename V>
__device__ __forceinline__ void store_wt(V* p, V v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// store_wt for a single filter's kernels (WT), a plain store otherwise: the server's batched
// kernels, whose grids cover many filters at once, ran 8% slower with device-scope stores.
template <bool WT, typename V>
__device__ __forceinline__ void store_out(V* p, V v) {
  if constexpr (WT)
    store_wt(p, v);
  else
    *p = v;
}

// ---- DPP wave primitives (gfx9 wave64) ---------------------------------------------
// Cross-lane moves on the VALU: row_shr:k within rows of 16, then row_bcast:15 / :31 to
// carry row totals across the wave.  Lanes without a source keep `old` (the identity).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_f(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v),
                                                    CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_d(double old, double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL,
                                             ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL,
                                             ROWS, 0xF, false);
  return __hiloint2double(hi, lo);
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114,
              kDppRowShr8 = 0x118, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
// Whole-wave shifts by one lane (gfx9): lane i reads lane i + 1 (shl) or i - 1 (shr); the
// lane without a source keeps `old`.
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;

// Inclusive sum scan over the 64 lanes (fixed association).
template <typename T>
__device__ __forceinline__ T wave_scan_dpp(T v);
template <>
__device__ __forceinline__ double wave_scan_dpp<double>(double v) {
  v += dpp_d<kDppRowShr1>(0.0, v);
  v += dpp_d<kDppRowShr2>(0.0, v);
  v += dpp_d<kDppRowShr4>(0.0, v);
  v += dpp_d<kDppRowShr8>(0.0, v);
  v += dpp_d<kDppRowBcast15, 0xA>(0.0, v);
  v += dpp_d<kDppRowBcast31, 0xC>(0.0, v);
  return v;
}
template <>
__device__ __forceinline__ float wave_scan_dpp<float>(float v) {
  v += dpp_f<kDppRowShr1>(0.f, v);
  v += dpp_f<kDppRowShr2>(0.f, v);
  v += dpp_f<kDppRowShr4>(0.f, v);
  v += dpp_f<kDppRowShr8>(0.f, v);
  v += dpp_f<kDppRowBcast15, 0xA>(0.f, v);
  v += dpp_f<kDppRowBcast31, 0xC>(0.f, v);
  return v;
}
// Inclusive max scan over the 64 lanes (exact).
__device__ __forceinline__ double wave_scan_max_dpp(double v) {
  const double ninf = -INFINITY;
  v = fmax(v, dpp_d<kDppRowShr1>(ninf, v));
  v = fmax(v, dpp_d<kDppRowShr2>(ninf, v));
  v = fmax(v, dpp_d<kDppRowShr4>(ninf, v));
  v = fmax(v, dpp_d<kDppRowShr8>(ninf, v));
  v = fmax(v, dpp_d<kDppRowBcast15, 0xA>(ninf, v));
  v = fmax(v, dpp_d<kDppRowBcast31, 0xC>(ninf, v));
  return v;
}
__device__ __forceinline__ float wave_scan_max_dpp(float v) {
  const float ninf = -INFINITY;
  v = fmaxf(v, dpp_f<kDppRowShr1>(ninf, v));
  v = fmaxf(v, dpp_f<kDppRowShr2>(ninf, v));
  v = fmaxf(v, dpp_f<kDppRowShr4>(ninf, v));
  v = fmaxf(v, dpp_f<kDppRowShr8>(ninf, v));
  v = fmaxf(v, dpp_f<kDppRowBcast15, 0xA>(ninf, v));
  v = fmaxf(v, dpp_f<kDppRowBcast31, 0xC>(ninf, v));
  return v;
}
// Inclusive max scan of int32 values over the 64 lanes.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ int dpp_i(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ int wave_scan_max_i32(int v) {
  constexpr int lo = -2147483647 - 1;
  v = max(v, dpp_i<kDppRowShr1>(lo, v));
  v = max(v, dpp_i<kDppRowShr2>(lo, v));
  v = max(v, dpp_i<kDppRowShr4>(lo, v));
  v = max(v, dpp_i<kDppRowShr8>(lo, v));
  v = max(v, dpp_i<kDppRowBcast15, 0xA>(lo, v));
  v = max(v, dpp_i<kDppRowBcast31, 0xC>(lo, v));
  return v;
}

// Orders one wave's LDS traffic: DS instructions of one wave complete in issue order, so a
// compiler-level fence at wavefront scope is enough between dependent phases (writes before
// the call are seen by reads after it from any lane of the same wave).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Lane 63's value (the scan total) broadcast to the wave.
__device__ __forceinline__ double lane63(double v) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}
__device__ __forceinline__ float lane63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// Wave reductions in a fixed order, result in every lane.
template <typename T>
__device__ __forceinline__ T wave_sum_dpp(T v) {
  return lane63(wave_scan_dpp<T>(v));
}
template <typename T>
__device__ __forceinline__ T wave_max_dpp(T v) {
  return lane63(wave_scan_max_dpp(v));
}

// (max, sum exp(v - max), sum exp(2 (v - max))) triples: log-sum-exp and ESS partials.
__device__ __forceinline__ float exp_t(float v) { return __expf(v); }
__device__ __forceinline__ double exp_t(double v) { return exp(v); }

template <typename T>
struct LseAcc {
  T m, s, s2;
};

template <typename T>
__device__ __forceinline__ LseAcc<T> lse_empty() {
  return LseAcc<T>{(T)-INFINITY, (T)0, (T)0};
}

template <typename T>
__device__ __forceinline__ LseAcc<T> lse_combine(LseAcc<T> a, LseAcc<T> b) {
  const T M = a.m > b.m ? a.m : b.m;
  if (M == (T)-INFINITY) return lse_empty<T>();
  const T ea = (a.m == (T)-INFINITY) ? (T)0 : exp_t(a.m - M);
  const T eb = (b.m == (T)-INFINITY) ? (T)0 : exp_t(b.m - M);
  return LseAcc<T>{M, a.s * ea + b.s * eb, a.s2 * ea * ea + b.s2 * eb * eb};
}

// Butterfly over the 64 lanes; every lane ends with the same (fixed-order) result.
template <typename T>
__device__ __forceinline__ LseAcc<T> wave_lse(LseAcc<T> a) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    LseAcc<T> o;
    o.m = wave_shfl_xor(a.m, m);
    o.s = wave_shfl_xor(a.s, m);
    o.s2 = wave_shfl_xor(a.s2, m);
    a = lse_combine(a, o);
  }
  return a;
}

}  // namespace ssm
