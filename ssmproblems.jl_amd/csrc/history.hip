// history.hip -- dense particle history of a bootstrap filter on the device.
//
// The device form of DenseParticleContainer (GF/src/containers.jl:328-418) filled as
// DenseAncestorCallback does (GF/src/callbacks.jl:53-80): the time-0 states at creation
// (PostInit), then after every step (PostUpdate) the particles' states, their normalised
// log-weights in FP64 and their ancestors (the identity after a step that kept its
// particles, preserve_sample, resamplers.jl:148-154).  Generation t (1-based, as the
// reference's time index) is stored as SoA planes states[t-1][d][Npad], logw[t-1][Npad]
// and anc[t-1][Npad], so a push is one coalesced copy of about 20 B per particle (D = 4,
// FP32) and the HBM footprint is (4 D + 12) B per particle and step: 2^20 particles over
// 1000 steps take 29 GB.
// Paths are read by walking the stored ancestors back from the last generation
// (get_ancestry, containers.jl:406-418), one thread per particle.
#include <algorithm>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "internal.h"

namespace ssm {

constexpr int kHistThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kHistThreads)
    k_hist_push(int N, int Npad, int D, const T* __restrict__ x, const T* __restrict__ logw,
                const int32_t* __restrict__ anc, const BfScalars* sc, T* st, double* lw,
                int32_t* an) {
  const int i = blockIdx.x * kHistThreads + threadIdx.x;
  if (i >= N) return;
  for (int d = 0; d < D; ++d) st[(size_t)d * Npad + i] = x[(size_t)d * Npad + i];
  // normalised log-weight l_i - lse in the filter's type, as ssm_bf_export reports it
  const T lse = (T)sc->lse;
  lw[i] = sc->typeless ? 0.0 : (double)(T)(logw[i] - lse);
  an[i] = sc->last_resampled ? anc[i] : i;
}

// The push of a recording history (also replayed from graphs): generation t (the step the
// finish kernel just counted, sc->step - 1) goes to slot t + *base, skipped outside
// [0, cap).
template <typename T>
__global__ void __launch_bounds__(kHistThreads)
    k_hist_push_rec(int N, int Npad, int D, const T* __restrict__ x, const T* __restrict__ logw,
                    const int32_t* __restrict__ anc, const BfScalars* sc,
                    const int32_t* __restrict__ base, int cap, T* st, double* lw, int32_t* an) {
  const int i = blockIdx.x * kHistThreads + threadIdx.x;
  const int slot = sc->step - 1 + *base;
  if (i >= N || slot < 0 || slot >= cap) return;
  const size_t o = (size_t)slot * Npad;
  for (int d = 0; d < D; ++d)
    st[o * D + (size_t)d * Npad + i] = x[(size_t)d * Npad + i];
  const T lse = (T)sc->lse;
  lw[o + i] = sc->typeless ? 0.0 : (double)(T)(logw[i] - lse);
  an[o + i] = sc->last_resampled ? anc[i] : i;
}

// Paths of particles i0 .. i0 + count - 1: out[(k (Tn + 1) + t) D + d], t = 0 .. Tn.
template <typename T>
__global__ void __launch_bounds__(kHistThreads)
    k_hist_paths(int count, int i0, int Npad, int D, int Tn, const T* __restrict__ x0,
                 const T* __restrict__ st, const int32_t* __restrict__ an, double* out) {
  const int k = blockIdx.x * kHistThreads + threadIdx.x;
  if (k >= count) return;
  const size_t plane = (size_t)D * Npad;
  int a = i0 + k;
  double* o = out + (size_t)k * (Tn + 1) * D;
  for (int t = Tn; t >= 1; --t) {
    const T* s = st + (size_t)(t - 1) * plane;
    for (int d = 0; d < D; ++d) o[(size_t)t * D + d] = (double)s[(size_t)d * Npad + a];
    a = an[(size_t)(t - 1) * Npad + a];
  }
  for (int d = 0; d < D; ++d) o[d] = (double)x0[(size_t)d * Npad + a];
}

}  // namespace ssm

using namespace ssm;

struct ssm_history {
  ssm_bf* bf = nullptr;  // null once the filter is destroyed
  ssm_ctx* ctx = nullptr;
  int N = 0, Npad = 0, D = 0, prec = SSM_F32;
  int T = 0;     // generations pushed
  int cap = 0;   // generations allocated
  void* x0 = nullptr;
  void* states = nullptr;
  double* logw = nullptr;
  int32_t* anc = nullptr;
  int32_t* base = nullptr;  // device slot base of a recording run
};

void ssm::history_detach_filter(ssm_history* h) { h->bf = nullptr; }

namespace {

size_t elem(const ssm_history* h) { return h->prec == SSM_F32 ? 4 : 8; }
size_t plane_bytes(const ssm_history* h) { return elem(h) * h->D * h->Npad; }

void hist_free(ssm_history* h) {
  void* ps[] = {h->x0, h->states, h->logw, h->anc, h->base};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

// Grows the generation capacity to `cap` (keeps the pushed generations).
int hist_grow(ssm_history* h, int cap) {
  ssm_ctx* ctx = h->ctx;
  void* st = nullptr;
  double* lw = nullptr;
  int32_t* an = nullptr;
  if (hipMalloc(&st, plane_bytes(h) * cap) != hipSuccess ||
      hipMalloc(&lw, sizeof(double) * h->Npad * cap) != hipSuccess ||
      hipMalloc(&an, sizeof(int32_t) * h->Npad * cap) != hipSuccess) {
    if (st) (void)hipFree(st);
    if (lw) (void)hipFree(lw);
    return set_error(ctx, SSM_ERR_OOM, "particle history: cannot hold %d generations of %d",
                     cap, h->N);
  }
  hipStream_t s = ctx->stream;
  if (h->T > 0) {
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(st, h->states, plane_bytes(h) * h->T,
                                      hipMemcpyDeviceToDevice, s));
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(lw, h->logw, sizeof(double) * h->Npad * h->T,
                                      hipMemcpyDeviceToDevice, s));
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(an, h->anc, sizeof(int32_t) * h->Npad * h->T,
                                      hipMemcpyDeviceToDevice, s));
  }
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (h->states) (void)hipFree(h->states);
  if (h->logw) (void)hipFree(h->logw);
  if (h->anc) (void)hipFree(h->anc);
  h->states = st;
  h->logw = lw;
  h->anc = an;
  h->cap = cap;
  return SSM_OK;
}

}  // namespace

extern "C" int32_t ssm_bf_history_create(ssm_bf* bf, int32_t capacity, ssm_history** out) {
  if (!bf || !out) return SSM_ERR_ARG;
  *out = nullptr;
  BfView v;
  int rc = bf_device_view(bf, &v);
  if (rc) return rc;
  ssm_ctx* ctx = v.ctx;
  if (capacity < 1)
    return set_error(ctx, SSM_ERR_ARG, "particle history: capacity must be >= 1 (got %d)",
                     capacity);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  ssm_history* h = new ssm_history();
  h->ctx = ctx;
  h->N = v.N;
  h->Npad = v.Npad;
  h->D = v.D;
  h->prec = v.prec;
  if (hipMalloc(&h->x0, plane_bytes(h)) != hipSuccess ||
      hipMalloc(&h->base, sizeof(int32_t)) != hipSuccess) {
    hist_free(h);
    delete h;
    return set_error(ctx, SSM_ERR_OOM, "particle history: time-0 states");
  }
  rc = hist_grow(h, capacity);
  if (rc == SSM_OK) rc = bf_attach_history(bf, h, true);
  if (rc) {
    hist_free(h);
    delete h;
    return rc;
  }
  h->bf = bf;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(h->x0, v.x, plane_bytes(h), hipMemcpyDeviceToDevice,
                                    ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *out = h;
  return SSM_OK;
}

namespace ssm {

int history_record_prepare(ssm_history* h, hipStream_t s, int t0, int T, bool* moved) {
  ssm_ctx* ctx = h->ctx;
  *moved = false;
  const int64_t need = (int64_t)h->T + T;
  if (need > INT32_MAX)
    return set_error(ctx, SSM_ERR_ARG, "particle history: 2^31 - 1 generations reached");
  if (need > h->cap) {
    int64_t cap2 = std::max<int64_t>(need, 2 * (int64_t)h->cap);
    const int rc = hist_grow(h, (int)std::min<int64_t>(cap2, INT32_MAX));
    if (rc) return rc;
    *moved = true;
  }
  const int32_t base = h->T - t0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(h->base, &base, sizeof base, hipMemcpyHostToDevice, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));  // `base` is a host local
  return SSM_OK;
}

void history_record_push(ssm_history* h, hipStream_t s, const BfView& v) {
  const dim3 g((h->N + kHistThreads - 1) / kHistThreads), b(kHistThreads);
  if (h->prec == SSM_F32)
    hipLaunchKernelGGL(k_hist_push_rec<float>, g, b, 0, s, h->N, h->Npad, h->D,
                       (const float*)v.x, (const float*)v.logw, v.anc, v.sc, h->base, h->cap,
                       (float*)h->states, h->logw, h->anc);
  else
    hipLaunchKernelGGL(k_hist_push_rec<double>, g, b, 0, s, h->N, h->Npad, h->D,
                       (const double*)v.x, (const double*)v.logw, v.anc, v.sc, h->base, h->cap,
                       (double*)h->states, h->logw, h->anc);
}

void history_record_commit(ssm_history* h, int T) { h->T += T; }

}  // namespace ssm

extern "C" int32_t ssm_bf_history_record(ssm_history* h, int32_t on) {
  if (!h) return SSM_ERR_ARG;
  ssm_ctx* ctx = h->ctx;
  if (!h->bf) return set_error(ctx, SSM_ERR_STATE, "particle history: its filter was destroyed");
  return bf_record_history(h->bf, h, on != 0);
}

extern "C" int32_t ssm_bf_history_backward_sample(ssm_history* h, const double* u,
                                                  int32_t* indices, double* trajectory) {
  if (!h || !u) return SSM_ERR_ARG;
  ssm_ctx* ctx = h->ctx;
  if (!h->bf) return set_error(ctx, SSM_ERR_STATE, "particle history: its filter was destroyed");
  if (h->T < 1) return set_error(ctx, SSM_ERR_STATE, "backward simulation: no generation stored");
  BsHistory hv{h->T, h->x0, h->states, h->logw};
  return bf_backward_sample(h->bf, hv, u, indices, trajectory, nullptr);
}

extern "C" int32_t ssm_bf_history_backward_sample_tv(ssm_history* h, const ssm_lg_model* dyn,
                                                     const double* u, int32_t* indices,
                                                     double* trajectory) {
  if (!h || !u || !dyn) return SSM_ERR_ARG;
  ssm_ctx* ctx = h->ctx;
  if (!h->bf) return set_error(ctx, SSM_ERR_STATE, "particle history: its filter was destroyed");
  if (h->T < 1) return set_error(ctx, SSM_ERR_STATE, "backward simulation: no generation stored");
  BsHistory hv{h->T, h->x0, h->states, h->logw};
  return bf_backward_sample(h->bf, hv, u, indices, trajectory, dyn);
}

extern "C" int32_t ssm_bf_history_reset(ssm_history* h) {
  if (!h) return SSM_ERR_ARG;
  ssm_ctx* ctx = h->ctx;
  if (!h->bf) return set_error(ctx, SSM_ERR_STATE, "particle history: its filter was destroyed");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  BfView v;
  const int rc = bf_device_view(h->bf, &v);
  if (rc) return rc;
  h->T = 0;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(h->x0, v.x, plane_bytes(h), hipMemcpyDeviceToDevice,
                                    ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return SSM_OK;
}

extern "C" int32_t ssm_bf_history_push(ssm_history* h) {
  if (!h) return SSM_ERR_ARG;
  ssm_ctx* ctx = h->ctx;
  if (!h->bf) return set_error(ctx, SSM_ERR_STATE, "particle history: its filter was destroyed");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  BfView v;
  int rc = bf_device_view(h->bf, &v);
  if (rc) return rc;
  if (h->T == h->cap) {
    const int64_t cap2 = std::min<int64_t>(2 * (int64_t)h->cap, INT32_MAX);
    if (cap2 == h->cap)
      return set_error(ctx, SSM_ERR_ARG, "particle history: 2^31 - 1 generations reached");
    rc = hist_grow(h, (int)cap2);
    if (rc) return rc;
  }
  const size_t t = (size_t)h->T;
  const dim3 g((h->N + kHistThreads - 1) / kHistThreads), b(kHistThreads);
  if (h->prec == SSM_F32) {
    const float* lwp = (const float*)v.logw;
    hipLaunchKernelGGL(k_hist_push<float>, g, b, 0, ctx->stream, h->N, h->Npad, h->D,
                       (const float*)v.x, lwp, v.anc, v.sc,
                       (float*)h->states + t * h->D * h->Npad, h->logw + t * h->Npad,
                       h->anc + t * h->Npad);
  } else {
    const double* lwp = (const double*)v.logw;
    hipLaunchKernelGGL(k_hist_push<double>, g, b, 0, ctx->stream, h->N, h->Npad, h->D,
                       (const double*)v.x, lwp, v.anc, v.sc,
                       (double*)h->states + t * h->D * h->Npad, h->logw + t * h->Npad,
                       h->anc + t * h->Npad);
  }
  SSM_HIP_CHECK(ctx, hipGetLastError());
  ++h->T;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_history_info(ssm_history* h, int32_t* T, int32_t* capacity) {
  if (!h) return SSM_ERR_ARG;
  if (T) *T = h->T;
  if (capacity) *capacity = h->cap;
  return SSM_OK;
}

extern "C" int32_t ssm_bf_history_ancestry(ssm_history* h, int32_t i0, int32_t count,
                                           double* paths) {
  if (!h || !paths) return SSM_ERR_ARG;
  ssm_ctx* ctx = h->ctx;
  if (i0 < 0 || count < 1 || (int64_t)i0 + count > h->N)
    return set_error(ctx, SSM_ERR_ARG, "particle history: particles %d..%lld outside 0..%d", i0,
                     (long long)i0 + count - 1, h->N - 1);
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)count * (h->T + 1) * h->D;
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * n) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "particle history: %zu doubles of paths", n);
  const dim3 g((count + kHistThreads - 1) / kHistThreads), b(kHistThreads);
  if (h->prec == SSM_F32)
    hipLaunchKernelGGL(k_hist_paths<float>, g, b, 0, s, count, i0, h->Npad, h->D, h->T,
                       (const float*)h->x0, (const float*)h->states, h->anc, d);
  else
    hipLaunchKernelGGL(k_hist_paths<double>, g, b, 0, s, count, i0, h->Npad, h->D, h->T,
                       (const double*)h->x0, (const double*)h->states, h->anc, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(paths, d, sizeof(double) * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess)
    return set_error(ctx, SSM_ERR_HIP, "particle history: path read-back failed: %s",
                     hipGetErrorString(e));
  return SSM_OK;
}

extern "C" int32_t ssm_bf_history_generation(ssm_history* h, int32_t t, double* x,
                                             double* logw, int32_t* ancestors) {
  if (!h) return SSM_ERR_ARG;
  ssm_ctx* ctx = h->ctx;
  if (t < 0 || t > h->T)
    return set_error(ctx, SSM_ERR_ARG, "particle history: time %d outside 0..%d", t, h->T);
  if (t == 0 && (logw || ancestors))
    return set_error(ctx, SSM_ERR_ARG, "particle history: time 0 holds states only");
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const char* src = (t == 0) ? (const char*)h->x0
                             : (const char*)h->states + plane_bytes(h) * (size_t)(t - 1);
  std::vector<char> xb;
  if (x) {
    xb.resize(plane_bytes(h));
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(xb.data(), src, xb.size(), hipMemcpyDeviceToHost, s));
  }
  if (logw)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(logw, h->logw + (size_t)(t - 1) * h->Npad,
                                      sizeof(double) * h->N, hipMemcpyDeviceToHost, s));
  if (ancestors)
    SSM_HIP_CHECK(ctx, hipMemcpyAsync(ancestors, h->anc + (size_t)(t - 1) * h->Npad,
                                      sizeof(int32_t) * h->N, hipMemcpyDeviceToHost, s));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (x) {
    for (int i = 0; i < h->N; ++i)
      for (int d = 0; d < h->D; ++d) {
        const size_t k = (size_t)d * h->Npad + i;
        x[(size_t)i * h->D + d] = (h->prec == SSM_F32) ? (double)((const float*)xb.data())[k]
                                                       : ((const double*)xb.data())[k];
      }
  }
  return SSM_OK;
}

extern "C" int32_t ssm_bf_history_destroy(ssm_history* h) {
  if (!h) return SSM_ERR_ARG;
  (void)hipSetDevice(h->ctx->device);
  (void)hipStreamSynchronize(h->ctx->stream);
  if (h->bf) (void)bf_attach_history(h->bf, h, false);
  hist_free(h);
  delete h;
  return SSM_OK;
}
