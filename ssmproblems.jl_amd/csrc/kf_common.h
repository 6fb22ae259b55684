// kf_common.h -- the batched Kalman kernels' shared parameter block, output block and
// input-layout helpers (kf.hip, kf_wg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

struct ssm_ctx;

namespace ssm {

struct KfBatchParams {
  int T;
  const int32_t* __restrict__ Tf;  // optional per-filter horizons (else T for all)
  int count;                      // filters in this class
  const int32_t* __restrict__ order;  // filter ids of this class, costliest first
  uint32_t* counter;              // work-queue head (zeroed before launch)
  const int32_t* __restrict__ Dx;
  const int32_t* __restrict__ Dy;
  const double* __restrict__ params;
  const int64_t* __restrict__ param_off;
  const double* __restrict__ ys;
  const int64_t* __restrict__ y_off;
  double* __restrict__ ll_out;
  int32_t* __restrict__ status;
  double* __restrict__ mu_out;     // optional final state (packed, offsets below)
  double* __restrict__ Sigma_out;
  const int64_t* __restrict__ mu_off;
  const int64_t* __restrict__ cov_off;
  // added to the diagonal of every filtered covariance (KalmanFilter's jitter; 0 = none)
  double jitter = 0.0;
  // kf_loglikelihood's layout (the TV kernels; uniform Dx, Dy): filter f starts from
  // mu0 + f Dx and Sigma0 + f Dx^2, and step t (0-based) uses parameter block
  // f * blocks + (blocks > 1 ? t : 0) of each array (column-major matrices)
  const double* __restrict__ tv_mu0 = nullptr;
  const double* __restrict__ tv_Sigma0 = nullptr;
  const double* __restrict__ tv_A = nullptr;
  const double* __restrict__ tv_b = nullptr;
  const double* __restrict__ tv_Q = nullptr;
  const double* __restrict__ tv_H = nullptr;
  const double* __restrict__ tv_c = nullptr;
  const double* __restrict__ tv_R = nullptr;
  int tv_blocks = 1;
  int fm = 0;  // the tiny kernels' inputs staged filter-minor (kf_at, kf_yat, kf_0at)
};

constexpr double kLog2Pi = 1.8378770664093454835606594728112;

// Indices into the tiny kernels' inputs: the caller's layout (parameter block f T + t, or f
// with one set a filter; mu0 and Sigma0 filter-major; ys at y_off[f]), or with p.fm the
// filter-minor staging of the same values, element e of step t of filter f at (t E + e) B + f,
// so the 64 lanes of a wave read 64 consecutive words.
__device__ __forceinline__ int64_t kf_at(const KfBatchParams& p, int f, int t, int64_t E, int e) {
  const int64_t tb = p.tv_blocks > 1 ? t : 0;
  return p.fm ? (tb * E + e) * (int64_t)p.count + f : ((int64_t)f * p.tv_blocks + tb) * E + e;
}
__device__ __forceinline__ int64_t kf_yat(const KfBatchParams& p, int f, int t, int dy, int q) {
  return p.fm ? ((int64_t)t * dy + q) * (int64_t)p.count + f
              : p.y_off[f] + (int64_t)t * dy + q;
}
__device__ __forceinline__ int64_t kf_0at(const KfBatchParams& p, int f, int64_t E, int e) {
  return p.fm ? (int64_t)e * p.count + f : (int64_t)f * E + e;
}


// Gradient outputs of kf_loglikelihood's reverse rule.
struct KfGradOut {
  double* mu0;    // B x Dx
  double* Sigma0; // B x Dx x Dx
  double *A, *b, *Q, *H, *c, *R;  // the parameter layout of the inputs
  double* cache;  // T x kC x B, kC = 4 N + 4 N^2 <= kGradCache
  int B;
};

// The workgroup kernels (kf_wg.hip) of the reverse rule and the square-root filter for
// filters beyond the lane kernels' 4 x 4, up to Dx = 64, Dy = 32: doubles of one filter's
// device workspace, and the launches over p's filters in chunks of `chunk` (g null: the
// square-root filter, writing mu_out and U_out).
int64_t kf_grad_wg_doubles(int dx, int dy, int T);
int64_t srkf_wg_doubles(int dx, int dy);
int launch_kf_wg(::ssm_ctx* ctx, hipStream_t s, const KfBatchParams& p, const KfGradOut* g,
                 double* mu_out, double* U_out, double* ws, int64_t per, int chunk);

}  // namespace ssm
