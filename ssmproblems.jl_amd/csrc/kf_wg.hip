// kf_wg.hip -- kf_loglikelihood's reverse rule and the square-root Kalman filter for filters
// of any shape up to Dx = 64, Dy = 32: one 256-thread workgroup per filter, FP64 throughout,
// each filter's matrices in a device workspace of its own (row-major, leading dimension the
// column count), every operation a workgroup-wide pass ended by a barrier.
//
// Reference (paths relative to the reference checkout):
//   reverse rule  GF/src/integrations/kalman_rrule.jl:22-107
//   backward      GF/src/algorithms/kalman_gradient.jl:103-263 (Parellier et al., eqs 8-14)
//   forward       kalman_gradient.jl:12-29 with kalman.jl:42-66 (Joseph form, Symmetric(:U))
//   SRKF          GF/src/algorithms/srkf.jl:93-176
//
// The gradient kernel keeps only the filtered state (mu, Sigma) of every step; the backward
// pass recomputes a step's predict and update from the state before it with the same
// device code, so its z, S, K and I - KH are the forward pass's bit for bit, and the
// workspace is (T + 1)(Dx + Dx^2) doubles plus one step's temporaries a filter.
#include <algorithm>
#include <cmath>

#include "device_common.h"
#include "internal.h"
#include "kf_common.h"

namespace ssm {
namespace {

constexpr int kWg = 256;
constexpr int kWgWaves = kWg / 64;

// ---- workgroup-wide FP64 linear algebra on row-major matrices in device memory --------

// Sum of one double per thread, in every thread (DPP wave sums, then the waves in order).
__device__ double wg_sum(double v, double* red) {
  v = wave_sum_dpp(v);
  __syncthreads();  // red is free
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kWgWaves; ++k) s += red[k];
  return s;
}

// C (m x n) = alpha op(A) op(B) [+ D], op(X) = X or X'; C may not alias A or B (D may be C).
__device__ void mm(double* C, int ldc, const double* A, int lda, bool tA, const double* B,
                   int ldb, bool tB, int m, int n, int k, const double* D = nullptr,
                   int ldd = 0, double alpha = 1.0) {
  for (int e = threadIdx.x; e < m * n; e += kWg) {
    const int i = e / n, j = e % n;
    double s = 0.0;
    for (int q = 0; q < k; ++q)
      s += (tA ? A[q * lda + i] : A[i * lda + q]) * (tB ? B[j * ldb + q] : B[q * ldb + j]);
    s *= alpha;
    if (D) s += D[i * ldd + j];
    C[i * ldc + j] = s;
  }
  __syncthreads();
}

// Lower Cholesky factor L (n x n, zero above the diagonal) of S.  A non-positive pivot makes
// it false and is replaced by 1, as the lane kernels do.
__device__ bool chol(double* L, const double* S, int n, int* flag) {
  if (threadIdx.x == 0) *flag = 1;
  for (int e = threadIdx.x; e < n * n; e += kWg)
    if (e % n > e / n) L[e] = 0.0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (threadIdx.x == 0) {
      double d = S[j * n + j];
      for (int k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
      if (!(d > 0.0)) {
        *flag = 0;
        d = 1.0;
      }
      L[j * n + j] = sqrt(d);
    }
    __syncthreads();
    const double ljj = L[j * n + j];
    for (int i = j + 1 + threadIdx.x; i < n; i += kWg) {
      double v = S[i * n + j];
      for (int k = 0; k < j; ++k) v -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = v / ljj;
    }
    __syncthreads();
  }
  const bool ok = *flag != 0;
  __syncthreads();
  return ok;
}

// X = (L L')^-1 V for the m columns of V (n x m, ld ldv; the identity when V is null), one
// column per thread by forward then backward substitution in scr (m x n).  tout: X holds
// the solutions as rows (X = ((L L')^-1 V)').
__device__ void cholsolve(double* X, int ldx, bool tout, const double* L, int n, const double* V,
                          int ldv, int m, double* scr) {
  for (int j = threadIdx.x; j < m; j += kWg) {
    double* y = scr + (size_t)j * n;
    for (int i = 0; i < n; ++i) {
      double s = V ? V[i * ldv + j] : (i == j ? 1.0 : 0.0);
      for (int k = 0; k < i; ++k) s -= L[i * n + k] * y[k];
      y[i] = s / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < n; ++k) s -= L[k * n + i] * y[k];
      y[i] = s / L[i * n + i];
    }
    for (int i = 0; i < n; ++i) {
      if (tout)
        X[j * ldx + i] = y[i];
      else
        X[i * ldx + j] = y[i];
    }
  }
  __syncthreads();
}

// R of a Householder QR of X (m x n, m >= n) in place: on return the top n rows hold R (upper
// triangular) and the rest zeros.  v: m doubles of scratch.
__device__ void qr_r(double* X, int m, int n, double* v, double* red) {
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = j + threadIdx.x; i < m; i += kWg) s += X[i * n + j] * X[i * n + j];
    s = wg_sum(s, red);
    if (!(s > 0.0)) continue;  // a zero column: nothing to reflect (uniform)
    const double nrm = sqrt(s), x0 = X[j * n + j];
    const double alpha = x0 >= 0.0 ? -nrm : nrm;  // v0 = x0 - alpha without cancellation
    const double v0 = x0 - alpha;
    const double vtv = s - x0 * x0 + v0 * v0;
    for (int i = j + threadIdx.x; i < m; i += kWg) v[i] = (i == j) ? v0 : X[i * n + j];
    __syncthreads();
    for (int k = j + 1 + threadIdx.x; k < n; k += kWg) {
      double d = 0.0;
      for (int i = j; i < m; ++i) d += v[i] * X[i * n + k];
      const double fac = 2.0 * d / vtv;
      for (int i = j; i < m; ++i) X[i * n + k] -= fac * v[i];
    }
    __syncthreads();
    for (int i = j + threadIdx.x; i < m; i += kWg) X[i * n + j] = (i == j) ? alpha : 0.0;
    __syncthreads();
  }
}

// _correct_cholesky_sign (srkf.jl:78-88): the first n rows of X (ld n) scaled by the sign of
// their diagonal entry (a zero diagonal zeroes the row).  The signs are read into sg (n
// doubles) before any row changes.
__device__ void sign_rows(double* X, int n, double* sg) {
  for (int i = threadIdx.x; i < n; i += kWg) {
    const double d = X[i * n + i];
    sg[i] = d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : 0.0);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += kWg) X[e] *= sg[e / n];
  __syncthreads();
}

// ---- one filter's step parameters and workspace ----------------------------------------

// Bump allocator over a filter's workspace (base null: only counts, for the host).
struct Bump {
  double* base;
  int64_t off = 0;
  __host__ __device__ double* take(int64_t n) {
    double* r = base ? base + off : nullptr;
    off += (n + 1) & ~(int64_t)1;  // 16-byte steps
    return r;
  }
};

// The step's parameters, row-major (A, Q: dx x dx; H: dy x dx; R: dy x dy).
struct StepPar {
  double *A, *b, *Q, *H, *c, *R, *y;
};

__device__ void load_step(const KfBatchParams& p, int f, int t, int dx, int dy, const StepPar& s) {
  const int64_t nxx = (int64_t)dx * dx, nyx = (int64_t)dy * dx, nyy = (int64_t)dy * dy;
  for (int e = threadIdx.x; e < dx * dx; e += kWg) {
    const int i = e / dx, j = e % dx;
    s.A[e] = p.tv_A[kf_at(p, f, t, nxx, i + j * dx)];
    s.Q[e] = p.tv_Q[kf_at(p, f, t, nxx, i + j * dx)];
  }
  for (int e = threadIdx.x; e < dy * dx; e += kWg) {
    const int i = e / dx, j = e % dx;
    s.H[e] = p.tv_H[kf_at(p, f, t, nyx, i + j * dy)];
  }
  for (int e = threadIdx.x; e < dy * dy; e += kWg) {
    const int i = e / dy, j = e % dy;
    s.R[e] = p.tv_R[kf_at(p, f, t, nyy, i + j * dy)];
  }
  for (int e = threadIdx.x; e < dx; e += kWg) s.b[e] = p.tv_b[kf_at(p, f, t, dx, e)];
  for (int e = threadIdx.x; e < dy; e += kWg) {
    s.c[e] = p.tv_c[kf_at(p, f, t, dy, e)];
    s.y[e] = p.ys[kf_yat(p, f, t, dy, e)];
  }
  __syncthreads();
}

// One Joseph-form step's intermediates.
struct Fwd {
  double *mp, *T1, *Sp, *z, *P, *S, *L, *K, *IKH, *KR, *mf, *scr;
};

// Predict and update of one step (kalman.jl:42-49, kalman_gradient.jl:12-29) from (mu, Sig)
// into (mu_new, Sig_new); *ll_inc (thread 0) the increment.  On return F.T1 holds
// (I - KH) Sigma_pred, F.P = H Sigma_pred.  False when S is not positive definite.
__device__ bool fwd_step(int dx, int dy, double jitter, const StepPar& s, const double* mu,
                         const double* Sig, const Fwd& F, double* mu_new, double* Sig_new,
                         double* ll_inc, int* flag) {
  for (int i = threadIdx.x; i < dx; i += kWg) {
    double v = s.b[i];
    for (int k = 0; k < dx; ++k) v += s.A[i * dx + k] * mu[k];
    F.mp[i] = v;
  }
  mm(F.T1, dx, s.A, dx, false, Sig, dx, false, dx, dx, dx);
  mm(F.Sp, dx, F.T1, dx, false, s.A, dx, true, dx, dx, dx, s.Q, dx);
  for (int q = threadIdx.x; q < dy; q += kWg) {
    double v = s.c[q];
    for (int k = 0; k < dx; ++k) v += s.H[q * dx + k] * F.mp[k];
    F.z[q] = s.y[q] - v;
  }
  mm(F.P, dx, s.H, dx, false, F.Sp, dx, false, dy, dx, dx);
  mm(F.S, dy, F.P, dx, false, s.H, dx, true, dy, dy, dx, s.R, dy);
  if (!chol(F.L, F.S, dy, flag)) return false;
  if (threadIdx.x == 0) {  // w = L^-1 z: the quadratic form and log-determinant
    double* w = F.scr;
    double qd = 0.0, ldet = 0.0;
    for (int i = 0; i < dy; ++i) {
      double v = F.z[i];
      for (int k = 0; k < i; ++k) v -= F.L[i * dy + k] * w[k];
      w[i] = v / F.L[i * dy + i];
      qd += w[i] * w[i];
      ldet += 2.0 * log(F.L[i * dy + i]);
    }
    *ll_inc = -0.5 * ((double)dy * kLog2Pi + ldet + qd);
  }
  __syncthreads();
  cholsolve(F.K, dy, true, F.L, dy, F.P, dx, dx, F.scr);  // K = (S^-1 P)'
  for (int e = threadIdx.x; e < dx * dx; e += kWg) {
    const int i = e / dx, j = e % dx;
    double u = (i == j) ? 1.0 : 0.0;
    for (int q = 0; q < dy; ++q) u -= F.K[i * dy + q] * s.H[q * dx + j];
    F.IKH[e] = u;
  }
  for (int e = threadIdx.x; e < dx * dy; e += kWg) {
    const int i = e / dy, j = e % dy;
    double r = 0.0;
    for (int q = 0; q < dy; ++q) r += F.K[i * dy + q] * s.R[q * dy + j];
    F.KR[e] = r;
  }
  for (int i = threadIdx.x; i < dx; i += kWg) {
    double v = F.mp[i];
    for (int q = 0; q < dy; ++q) v += F.K[i * dy + q] * F.z[q];
    F.mf[i] = v;
  }
  __syncthreads();
  mm(F.T1, dx, F.IKH, dx, false, F.Sp, dx, false, dx, dx, dx);
  // the upper triangle of T1 (I-KH)' + KR K', mirrored (Symmetric(:U)), plus jitter
  for (int e = threadIdx.x; e < dx * dx; e += kWg) {
    const int i = e / dx, j = e % dx;
    if (j < i) continue;
    double v = 0.0;
    for (int k = 0; k < dx; ++k) v += F.T1[i * dx + k] * F.IKH[j * dx + k];
    for (int q = 0; q < dy; ++q) v += F.KR[i * dy + q] * F.K[j * dy + q];
    if (i == j) v += jitter;
    Sig_new[i * dx + j] = v;
    Sig_new[j * dx + i] = v;
  }
  for (int i = threadIdx.x; i < dx; i += kWg) mu_new[i] = F.mf[i];
  __syncthreads();
  return true;
}

}  // namespace

// ---- kf_loglikelihood's reverse rule ---------------------------------------------------
// The workspace of one filter: the cached states (mu then Sigma, T + 1 of them), the step's
// parameters, the forward step's intermediates and the backward pass's.
struct GradWs {
  double* states;
  StepPar s;
  Fwd F;
  double *mu_x, *Sig_x;  // the recomputed filtered state (unused by the backward pass)
  double *Siz, *Sinv, *Ktd, *KtdS, *T2, *SinvP, *rowv, *KtdSK, *HtSiz, *dmp, *LR, *Riz, *HtRiz,
      *inner, *T3, *SinvH, *HtSinvH, *dSp, *T4, *dmu, *dS;
};

// Lays out one filter's workspace at base (null: sizes only); returns its doubles.
__host__ __device__ int64_t grad_ws_layout(int dx, int dy, int T, double* base, GradWs* w) {
  const int64_t nxx = (int64_t)dx * dx, nyx = (int64_t)dy * dx, nyy = (int64_t)dy * dy;
  const int64_t mr = dx > dy ? dx : dy;
  Bump bp{base};
  GradWs g;
  g.states = bp.take((T + 1) * (dx + nxx));
  g.s = StepPar{bp.take(nxx), bp.take(dx), bp.take(nxx), bp.take(nyx), bp.take(dy),
                bp.take(nyy), bp.take(dy)};
  g.F = Fwd{bp.take(dx),  bp.take(nxx), bp.take(nxx), bp.take(dy),  bp.take(nyx),
            bp.take(nyy), bp.take(nyy), bp.take(nyx), bp.take(nxx), bp.take(nyx),
            bp.take(dx),  bp.take(mr * mr)};
  g.mu_x = bp.take(dx);
  g.Sig_x = bp.take(nxx);
  g.Siz = bp.take(dy);
  g.Sinv = bp.take(nyy);
  g.Ktd = bp.take(dy);
  g.KtdS = bp.take(nyx);
  g.T2 = bp.take(nyx);
  g.SinvP = bp.take(nyx);
  g.rowv = bp.take(2 * (int64_t)dx);  // Siz' P and dmu' Pf
  g.KtdSK = bp.take(nyy);
  g.HtSiz = bp.take(dx);
  g.dmp = bp.take(dx);
  g.LR = bp.take(nyy);
  g.Riz = bp.take(dy);
  g.HtRiz = bp.take(dx);
  g.inner = bp.take(nxx);
  g.T3 = bp.take(nxx);
  g.SinvH = bp.take(nyx);
  g.HtSinvH = bp.take(nxx);
  g.dSp = bp.take(nxx);
  g.T4 = bp.take(nxx);
  g.dmu = bp.take(dx);
  g.dS = bp.take(nxx);
  if (w) *w = g;
  return bp.off;
}

// Filter f = f0 + blockIdx.x; ws_all: this launch's workspace, ws_per doubles a filter.
__global__ void __launch_bounds__(kWg)
    k_kf_grad_wg(KfBatchParams p, KfGradOut g, double* ws_all, int64_t ws_per, int f0) {
  const int f = f0 + blockIdx.x;
  if (f >= p.count) return;
  __shared__ double red[kWgWaves];
  __shared__ double sh_inc;
  __shared__ int flag;
  (void)red;
  const int dx = p.Dx[f], dy = p.Dy[f];
  const int Tn = p.T;
  const int64_t nxx = (int64_t)dx * dx, nyx = (int64_t)dy * dx, nyy = (int64_t)dy * dy;
  const int64_t st = dx + nxx;  // a cached state: mu then Sigma
  GradWs W;
  grad_ws_layout(dx, dy, Tn, ws_all + (int64_t)blockIdx.x * ws_per, &W);
  double* const states = W.states;
  const StepPar& s = W.s;
  const Fwd& F = W.F;
  // ---- forward (kalman_rrule.jl:27-50) ----
  for (int e = threadIdx.x; e < dx; e += kWg) states[e] = p.tv_mu0[kf_0at(p, f, dx, e)];
  for (int e = threadIdx.x; e < dx * dx; e += kWg) {
    const int i = e / dx, j = e % dx;
    states[dx + e] = p.tv_Sigma0[kf_0at(p, f, nxx, i + j * dx)];
  }
  __syncthreads();
  double ll = 0.0;
  bool ok = true;
  for (int t = 0; t < Tn && ok; ++t) {
    load_step(p, f, t, dx, dy, s);
    double* cur = states + (int64_t)t * st;
    double* nxt = cur + st;
    ok = fwd_step(dx, dy, p.jitter, s, cur, cur + dx, F, nxt, nxt + dx, &sh_inc, &flag);
    if (ok && threadIdx.x == 0) ll += sh_inc;
  }
  if (threadIdx.x == 0) {
    p.ll_out[f] = ok ? ll : NAN;
    p.status[f] = ok ? SSM_OK : SSM_ERR_NOT_PD;
  }
  if (!ok) return;
  // ---- backward (kalman_rrule.jl:63-103), NLL convention; the outputs are negated ----
  double *Siz = W.Siz, *Sinv = W.Sinv, *Ktd = W.Ktd, *KtdS = W.KtdS, *T2 = W.T2,
         *SinvP = W.SinvP, *rowv = W.rowv, *KtdSK = W.KtdSK, *HtSiz = W.HtSiz, *dmp = W.dmp,
         *LR = W.LR, *Riz = W.Riz, *HtRiz = W.HtRiz, *inner = W.inner, *T3 = W.T3,
         *SinvH = W.SinvH, *HtSinvH = W.HtSinvH, *dSp = W.dSp, *T4 = W.T4, *dmu = W.dmu,
         *dS = W.dS;
  const bool acc = p.tv_blocks <= 1;  // one parameter set for all steps: sum the steps
  auto blk = [&](int t) -> int64_t { return (int64_t)f * p.tv_blocks + (p.tv_blocks > 1 ? t : 0); };
  // out[off + i + j r] (column-major) = (acc ? out : 0) - X[i][j]; each entry has one owner
  auto put = [&](double* out, int64_t off, int r, int cc, const double* X) {
    for (int e = threadIdx.x; e < r * cc; e += kWg) {
      const int i = e / cc, j = e % cc;
      double& o = out[off + i + (int64_t)j * r];
      o = (acc ? o : 0.0) - X[e];
    }
  };
  if (acc) {
    const int64_t k = blk(0);
    for (int e = threadIdx.x; e < nxx; e += kWg) {
      g.A[k * nxx + e] = 0.0;
      g.Q[k * nxx + e] = 0.0;
    }
    for (int e = threadIdx.x; e < nyx; e += kWg) g.H[k * nyx + e] = 0.0;
    for (int e = threadIdx.x; e < nyy; e += kWg) g.R[k * nyy + e] = 0.0;
    for (int e = threadIdx.x; e < dx; e += kWg) g.b[k * dx + e] = 0.0;
    for (int e = threadIdx.x; e < dy; e += kWg) g.c[k * dy + e] = 0.0;
  }
  for (int e = threadIdx.x; e < dx; e += kWg) dmu[e] = 0.0;
  for (int e = threadIdx.x; e < nxx; e += kWg) dS[e] = 0.0;
  __syncthreads();
  for (int t = Tn - 1; t >= 0; --t) {
    const int64_t k = blk(t);
    load_step(p, f, t, dx, dy, s);
    const double* mprev = states + (int64_t)t * st;
    const double* Sprev = mprev + dx;
    fwd_step(dx, dy, p.jitter, s, mprev, Sprev, F, W.mu_x, W.Sig_x, &sh_inc, &flag);
    // S^-1 z and S^-1
    cholsolve(Siz, 1, false, F.L, dy, F.z, 1, 1, F.scr);
    cholsolve(Sinv, dy, false, F.L, dy, nullptr, 0, dy, F.scr);
    for (int q = threadIdx.x; q < dy; q += kWg) {  // K' dmu
      double v = 0.0;
      for (int i = 0; i < dx; ++i) v += F.K[i * dy + q] * dmu[i];
      Ktd[q] = v;
    }
    __syncthreads();
    // gradient_c (kalman_gradient.jl:256-263): -(S^-1 z) - K' dmu
    for (int q = threadIdx.x; q < dy; q += kWg) {
      double& o = g.c[k * dy + q];
      o = (acc ? o : 0.0) - (-Siz[q] - Ktd[q]);
    }
    // gradient_H (kalman_gradient.jl:216-246); P_filt = (I-KH) Sigma_pred = F.T1, H Sigma_pred
    // = F.P
    mm(KtdS, dx, F.K, dy, true, dS, dx, false, dy, dx, dx);
    mm(T2, dx, KtdS, dx, false, F.T1, dx, false, dy, dx, dx);
    mm(SinvP, dx, Sinv, dy, false, F.P, dx, false, dy, dx, dy);
    for (int j = threadIdx.x; j < dx; j += kWg) {
      double a = 0.0, c2 = 0.0;
      for (int q = 0; q < dy; ++q) a += Siz[q] * F.P[q * dx + j];
      for (int i = 0; i < dx; ++i) c2 += dmu[i] * F.T1[i * dx + j];
      rowv[j] = a;
      rowv[dx + j] = c2;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < dy * dx; e += kWg) {
      const int q = e / dx, j = e % dx;
      const double v = SinvP[e] - Siz[q] * F.mp[j] - Siz[q] * rowv[j] + Siz[q] * rowv[dx + j] -
                       Ktd[q] * F.mf[j] - 2.0 * T2[e];
      double& o = g.H[k * nyx + q + (int64_t)j * dy];
      o = (acc ? o : 0.0) - v;
    }
    // gradient_R (kalman_gradient.jl:160-174)
    mm(KtdSK, dy, KtdS, dx, false, F.K, dy, false, dy, dy, dx);
    for (int e = threadIdx.x; e < dy * dy; e += kWg) {
      const int q = e / dy, r = e % dy;
      const double v = KtdSK[e] - 0.5 * (Ktd[q] * Siz[r] + Siz[q] * Ktd[r]) +
                       0.5 * (Sinv[e] - Siz[q] * Siz[r]);
      double& o = g.R[k * nyy + q + (int64_t)r * dy];
      o = (acc ? o : 0.0) - v;
    }
    // backward_gradient_update (kalman_gradient.jl:103-124)
    for (int i = threadIdx.x; i < dx; i += kWg) {
      double a = 0.0, u = 0.0;
      for (int q = 0; q < dy; ++q) a += s.H[q * dx + i] * Siz[q];
      for (int q = 0; q < dx; ++q) u += F.IKH[q * dx + i] * dmu[q];
      HtSiz[i] = a;
      dmp[i] = u - a;
    }
    chol(LR, s.R, dy, &flag);
    cholsolve(Riz, 1, false, LR, dy, F.z, 1, 1, F.scr);
    for (int i = threadIdx.x; i < dx; i += kWg) {
      double a = 0.0;
      for (int q = 0; q < dy; ++q) a += s.H[q * dx + i] * Riz[q];
      HtRiz[i] = a;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nxx; e += kWg) {
      const int i = e / dx, j = e % dx;
      inner[e] = dS[e] + 0.5 * (dmu[i] * HtRiz[j] + HtRiz[i] * dmu[j]);
    }
    __syncthreads();
    mm(T3, dx, F.IKH, dx, true, inner, dx, false, dx, dx, dx);   // IKH' inner
    mm(dSp, dx, T3, dx, false, F.IKH, dx, false, dx, dx, dx);    // IKH' inner IKH
    mm(SinvH, dx, Sinv, dy, false, s.H, dx, false, dy, dx, dy);
    mm(HtSinvH, dx, s.H, dx, true, SinvH, dx, false, dx, dx, dy);
    for (int e = threadIdx.x; e < nxx; e += kWg) {
      const int i = e / dx, j = e % dx;
      dSp[e] += 0.5 * (HtSinvH[e] - HtSiz[i] * HtSiz[j]);
    }
    __syncthreads();
    // dynamics parameters (kalman_gradient.jl:150-152, :196-208)
    for (int i = threadIdx.x; i < dx; i += kWg) {
      double& o = g.b[k * dx + i];
      o = (acc ? o : 0.0) - dmp[i];
    }
    put(g.Q, k * nxx, dx, dx, dSp);
    mm(T4, dx, dSp, dx, false, s.A, dx, false, dx, dx, dx);     // dSp A
    mm(T3, dx, T4, dx, false, Sprev, dx, false, dx, dx, dx);    // dSp A Sigma_prev
    for (int e = threadIdx.x; e < nxx; e += kWg) {
      const int i = e / dx, j = e % dx;
      const double v = dmp[i] * mprev[j] + 2.0 * T3[e];
      double& o = g.A[k * nxx + i + (int64_t)j * dx];
      o = (acc ? o : 0.0) - v;
    }
    // backward_gradient_predict (kalman_gradient.jl:139-143)
    for (int i = threadIdx.x; i < dx; i += kWg) {
      double a = 0.0;
      for (int q = 0; q < dx; ++q) a += s.A[q * dx + i] * dmp[q];
      dmu[i] = a;
    }
    mm(T4, dx, s.A, dx, true, dSp, dx, false, dx, dx, dx);      // A' dSp
    mm(dS, dx, T4, dx, false, s.A, dx, false, dx, dx, dx);      // A' dSp A
  }
  for (int i = threadIdx.x; i < dx; i += kWg) g.mu0[(int64_t)f * dx + i] = -dmu[i];
  for (int e = threadIdx.x; e < nxx; e += kWg) {
    const int i = e / dx, j = e % dx;
    g.Sigma0[(int64_t)f * nxx + i + (int64_t)j * dx] = -dS[e];
  }
}

// Doubles of one filter's workspace in k_kf_grad_wg.
int64_t kf_grad_wg_doubles(int dx, int dy, int T) { return grad_ws_layout(dx, dy, T, nullptr, nullptr); }

// ---- square-root Kalman filter (srkf.jl:93-176) -----------------------------------------
// The covariance is carried as its upper Cholesky factor U (Sigma = U' U).  Predict: the R
// of a QR factorisation of [U_Q; U A'] (2 Dx x Dx), rows made to have a positive diagonal.
// Update: the R of [[U_R, 0], [U H', U]] ((Dy + Dx) square), sign-corrected, gives U_S, the
// block X = R[0:Dy, Dy:] (so P H' = X') and the filtered factor R[Dy:, Dy:];
// w = U_S'^-1 z, mu += X' w and ll += -(w'w + 2 sum log diag U_S + Dy log 2 pi) / 2; the
// jitter is added to the factor's diagonal, as srkf_update does.
struct SrWs {
  StepPar s;
  double *mu, *U, *L, *UQ, *UR, *M1, *M2, *v, *w, *S0;
};

__host__ __device__ int64_t srkf_ws_layout(int dx, int dy, double* base, SrWs* w) {
  const int64_t nxx = (int64_t)dx * dx, nyx = (int64_t)dy * dx, nyy = (int64_t)dy * dy;
  const int64_t n2 = (int64_t)dx + dy, mr = dx > dy ? dx : dy;
  Bump bp{base};
  SrWs g;
  g.s = StepPar{bp.take(nxx), bp.take(dx), bp.take(nxx), bp.take(nyx), bp.take(dy),
                bp.take(nyy), bp.take(dy)};
  g.mu = bp.take(dx);
  g.U = bp.take(nxx);
  g.L = bp.take(mr * mr);
  g.UQ = bp.take(nxx);
  g.UR = bp.take(nyy);
  g.M1 = bp.take(2 * nxx);
  g.M2 = bp.take(n2 * n2);
  g.v = bp.take(2 * mr + n2);
  g.w = bp.take(dy);
  g.S0 = bp.take(nxx);
  if (w) *w = g;
  return bp.off;
}

__global__ void __launch_bounds__(kWg)
    k_srkf_wg(KfBatchParams p, double* mu_out, double* U_out, double* ws_all, int64_t ws_per,
              int f0) {
  const int f = f0 + blockIdx.x;
  if (f >= p.count) return;
  __shared__ double red[kWgWaves];
  __shared__ int flag;
  const int dx = p.Dx[f], dy = p.Dy[f];
  const int n2 = dx + dy;
  const int64_t nxx = (int64_t)dx * dx;
  SrWs W;
  srkf_ws_layout(dx, dy, ws_all + (int64_t)blockIdx.x * ws_per, &W);
  const StepPar& s = W.s;
  double *mu = W.mu, *U = W.U, *L = W.L, *UQ = W.UQ, *UR = W.UR, *M1 = W.M1, *M2 = W.M2,
         *v = W.v, *w = W.w, *S0 = W.S0;
  // U0 = chol(Sigma0).U
  for (int e = threadIdx.x; e < dx; e += kWg) mu[e] = p.tv_mu0[kf_0at(p, f, dx, e)];
  for (int e = threadIdx.x; e < dx * dx; e += kWg) {
    const int i = e / dx, j = e % dx;
    S0[e] = p.tv_Sigma0[kf_0at(p, f, nxx, i + j * dx)];
  }
  __syncthreads();
  bool ok = chol(L, S0, dx, &flag);
  for (int e = threadIdx.x; e < dx * dx; e += kWg) U[e] = L[(e % dx) * dx + e / dx];
  __syncthreads();
  double ll = 0.0;
  for (int t = 0; t < p.T && ok; ++t) {
    load_step(p, f, t, dx, dy, s);
    ok = chol(L, s.Q, dx, &flag);
    for (int e = threadIdx.x; e < dx * dx; e += kWg) UQ[e] = L[(e % dx) * dx + e / dx];
    __syncthreads();
    ok = chol(L, s.R, dy, &flag) && ok;
    for (int e = threadIdx.x; e < dy * dy; e += kWg) UR[e] = L[(e % dy) * dy + e / dy];
    // predict: mu = A mu + b (into w's neighbour v, then back), M1 = [U_Q; U A']
    for (int i = threadIdx.x; i < dx; i += kWg) {
      double a = s.b[i];
      for (int k = 0; k < dx; ++k) a += s.A[i * dx + k] * mu[k];
      v[i] = a;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < dx; i += kWg) mu[i] = v[i];
    for (int e = threadIdx.x; e < dx * dx; e += kWg) M1[e] = UQ[e];
    mm(M1 + nxx, dx, U, dx, false, s.A, dx, true, dx, dx, dx);
    qr_r(M1, 2 * dx, dx, v, red);
    sign_rows(M1, dx, v);
    // update: M2 = [[U_R, 0], [U H', U]]
    for (int e = threadIdx.x; e < n2 * n2; e += kWg) {
      const int i = e / n2, j = e % n2;
      double a = 0.0;
      if (i < dy) {
        a = (j < dy) ? UR[i * dy + j] : 0.0;
      } else if (j >= dy) {
        a = M1[(i - dy) * dx + (j - dy)];
      } else {
        for (int k = 0; k < dx; ++k) a += M1[(i - dy) * dx + k] * s.H[j * dx + k];
      }
      M2[e] = a;
    }
    __syncthreads();
    qr_r(M2, n2, n2, v, red);
    sign_rows(M2, n2, v);
    if (threadIdx.x == 0) {
      // z = y - H mu - c; w = U_S'^-1 z (U_S' lower); ll
      double qd = 0.0, ldet = 0.0;
      for (int i = 0; i < dy; ++i) {
        double zi = s.y[i] - s.c[i];
        for (int k = 0; k < dx; ++k) zi -= s.H[i * dx + k] * mu[k];
        for (int k = 0; k < i; ++k) zi -= M2[k * n2 + i] * w[k];
        w[i] = zi / M2[i * n2 + i];
        qd += w[i] * w[i];
        ldet += 2.0 * log(M2[i * n2 + i]);
      }
      ll += -0.5 * (qd + ldet + (double)dy * kLog2Pi);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < dx; i += kWg) {
      double a = mu[i];
      for (int q = 0; q < dy; ++q) a += M2[q * n2 + dy + i] * w[q];
      v[i] = a;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < dx; i += kWg) mu[i] = v[i];
    for (int e = threadIdx.x; e < dx * dx; e += kWg) {
      const int i = e / dx, j = e % dx;
      U[e] = M2[(dy + i) * n2 + dy + j] + ((i == j) ? p.jitter : 0.0);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    p.ll_out[f] = ok ? ll : NAN;
    p.status[f] = ok ? SSM_OK : SSM_ERR_NOT_PD;
  }
  if (mu_out)
    for (int i = threadIdx.x; i < dx; i += kWg) mu_out[(int64_t)f * dx + i] = mu[i];
  if (U_out)
    for (int e = threadIdx.x; e < dx * dx; e += kWg) {
      const int i = e / dx, j = e % dx;
      U_out[(int64_t)f * nxx + i + (int64_t)j * dx] = U[e];
    }
}

// Doubles of one filter's workspace in k_srkf_wg.
int64_t srkf_wg_doubles(int dx, int dy) { return srkf_ws_layout(dx, dy, nullptr, nullptr); }

// Enqueues the workgroup kernels over the B filters of p in chunks of `chunk` filters, each
// chunk's workspace at ws (chunk * per doubles).
int launch_kf_wg(ssm_ctx* ctx, hipStream_t s, const KfBatchParams& p, const KfGradOut* g,
                 double* mu_out, double* U_out, double* ws, int64_t per, int chunk) {
  for (int f0 = 0; f0 < p.count; f0 += chunk) {
    const int n = std::min(chunk, p.count - f0);
    if (g)
      hipLaunchKernelGGL(k_kf_grad_wg, dim3(n), dim3(kWg), 0, s, p, *g, ws, per, f0);
    else
      hipLaunchKernelGGL(k_srkf_wg, dim3(n), dim3(kWg), 0, s, p, mu_out, U_out, ws, per, f0);
    SSM_HIP_CHECK(ctx, hipGetLastError());
  }
  return SSM_OK;
}

}  // namespace ssm
