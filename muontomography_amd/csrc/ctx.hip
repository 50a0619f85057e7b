// Context: device selection, stream, last error.
#include <string.h>

#include "muto_internal.hpp"

extern "C" int muto_abi_version(void) { return MUTO_ABI_VERSION; }

extern "C" int muto_ctx_create(int device, muto_ctx** out) {
  if (!out) return MUTO_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MUTO_ERR_NODEV;
  if (device < 0 || device >= n) return MUTO_ERR_NODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MUTO_ERR_HIP;
  // the code objects are built for gfx950 only
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return MUTO_ERR_NODEV;
  if (hipSetDevice(device) != hipSuccess) return MUTO_ERR_HIP;
  // the no-alloc entry point's scratch address on this device (its status is
  // kept and reported by that entry point; nothing else depends on it)
  (void)muto::noalloc_scratch_init(device);
  auto* c = new muto_ctx();
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MUTO_ERR_HIP;
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MUTO_ERR_HIP;
  }
  *out = c;
  return MUTO_OK;
}

extern "C" int muto_ctx_destroy(muto_ctx* ctx) {
  if (!ctx) return MUTO_OK;
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  for (void* w : ctx->ws)
    if (w) (void)hipFree(w);
  delete ctx;
  return MUTO_OK;
}

extern "C" const char* muto_last_error(const muto_ctx* ctx) {
  if (!ctx) return "null context";
  return ctx->err.c_str();
}

extern "C" int muto_ctx_kernel_ms(muto_ctx* ctx, double* ms) {
  if (!ctx || !ms) return MUTO_ERR_ARG;
  *ms = 0.0;
  if (!ctx->ev_valid) {
    ctx->set_error("muto_ctx_kernel_ms: no timed call on this context");
    return MUTO_ERR_ARG;
  }
  float t = 0.f;
  MUTO_CHECK_HIP(ctx, hipEventSynchronize(ctx->ev1));
  MUTO_CHECK_HIP(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  *ms = t;
  return MUTO_OK;
}
