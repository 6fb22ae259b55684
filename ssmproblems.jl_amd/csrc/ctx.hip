// ctx.hip -- library context, error reporting and device-buffer helpers.
#include "internal.h"

namespace ssm {

int set_error(ssm_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

int ensure_device(ssm_ctx* ctx, void** p, size_t* cap, size_t bytes) {
  if (*cap >= bytes && *p) return SSM_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, bytes) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "cannot allocate %zu bytes of device memory", bytes);
  *cap = bytes;
  return SSM_OK;
}

}  // namespace ssm

using namespace ssm;

extern "C" int32_t ssm_version(void) { return 10000; }

extern "C" int32_t ssm_ctx_create(int32_t device, uint64_t hbm_budget_bytes, ssm_ctx** out) {
  if (!out) return SSM_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SSM_ERR_NO_DEVICE;
  if (device < 0 || device >= count) return SSM_ERR_ARG;
  ssm_ctx* c = new ssm_ctx();
  c->device = device;
  c->hbm_budget = hbm_budget_bytes;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SSM_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
  // the no-allocation resampler's scratch address on this device (the device is current)
  if (noalloc_setup(c) != SSM_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return SSM_ERR_HIP;
  }
  *out = c;
  return SSM_OK;
}

extern "C" int32_t ssm_ctx_destroy(ssm_ctx* ctx) {
  if (!ctx) return SSM_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->rs_scratch) (void)hipFree(ctx->rs_scratch);
  if (ctx->kfll) (void)hipFree(ctx->kfll);
  if (ctx->kfwg) (void)hipFree(ctx->kfwg);
  for (hipEvent_t e : ctx->kfll_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->rs_w) (void)hipFree(ctx->rs_w);
  if (ctx->rs_u) (void)hipFree(ctx->rs_u);
  if (ctx->rs_a) (void)hipFree(ctx->rs_a);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SSM_OK;
}

extern "C" const char* ssm_last_error(const ssm_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int32_t ssm_ctx_synchronize(ssm_ctx* ctx) {
  if (!ctx) return SSM_ERR_ARG;
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return SSM_OK;
}

extern "C" int32_t ssm_ctx_device_info(ssm_ctx* ctx, char* name, int32_t name_len,
                                       int32_t* cu_count) {
  if (!ctx) return SSM_ERR_ARG;
  hipDeviceProp_t prop;
  SSM_HIP_CHECK(ctx, hipGetDeviceProperties(&prop, ctx->device));
  if (name && name_len > 0) {
    snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  return SSM_OK;
}

extern "C" int32_t ssm_dev_alloc(ssm_ctx* ctx, uint64_t bytes, void** out) {
  if (!ctx || !out) return SSM_ERR_ARG;
  SSM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess)
    return set_error(ctx, SSM_ERR_OOM, "ssm_dev_alloc: %llu bytes", (unsigned long long)bytes);
  return SSM_OK;
}

extern "C" int32_t ssm_dev_free(ssm_ctx* ctx, void* p) {
  if (!ctx) return SSM_ERR_ARG;
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  SSM_HIP_CHECK(ctx, hipFree(p));
  return SSM_OK;
}

extern "C" int32_t ssm_memcpy_h2d(ssm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx) return SSM_ERR_ARG;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return SSM_OK;
}

extern "C" int32_t ssm_memcpy_d2h(ssm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx) return SSM_ERR_ARG;
  SSM_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SSM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return SSM_OK;
}
