// Internal declarations shared by the HIP translation units of libmuto.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "../../include/muto.h"

#define MUTO_CHECK_HIP(ctx, expr)                                          \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      (ctx)->set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
      return e_ == hipErrorOutOfMemory ? MUTO_ERR_NOMEM : MUTO_ERR_HIP;    \
    }                                                                      \
  } while (0)

struct muto_ctx {
  int device = 0;
  int n_cu = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // events around the kernels of the last one-shot call (Siddon, PoCA),
  // read by muto_ctx_kernel_ms
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool ev_valid = false, ev_begun = false;
  // device workspaces of the one-shot image calls (PoCA, PoCut, SLR), kept
  // for the next call and grown when one needs more; freed with the context
  void* ws[2] = {nullptr, nullptr};
  size_t ws_bytes[2] = {0, 0};
  hipError_t workspace(int k, size_t bytes, void** out) {
    if (bytes > ws_bytes[k]) {
      if (ws[k]) {
        const hipError_t e = hipFree(ws[k]);
        ws[k] = nullptr;
        ws_bytes[k] = 0;
        if (e != hipSuccess) return e;
      }
      const size_t b = bytes + bytes / 4;
      const hipError_t e = hipMalloc(&ws[k], b);
      if (e != hipSuccess) {
        ws[k] = nullptr;
        return e;
      }
      ws_bytes[k] = b;
    }
    *out = ws[k];
    return hipSuccess;
  }
  void set_error(const std::string& m) { err = m; }
  void mark_begin() {
    ev_valid = false;
    ev_begun = hipEventRecord(ev0, stream) == hipSuccess;
  }
  void mark_end() { ev_valid = ev_begun && hipEventRecord(ev1, stream) == hipSuccess; }
};

namespace muto {

// Grid as the kernels see it.
struct GridD {
  uint32_t n[3];
  double mn[3];
  double d[3];
};

inline GridD to_dev(const muto_grid& g) {
  GridD r;
  r.n[0] = g.nx; r.n[1] = g.ny; r.n[2] = g.nz;
  r.mn[0] = g.x_min; r.mn[1] = g.y_min; r.mn[2] = g.z_min;
  r.d[0] = g.dx; r.d[1] = g.dy; r.d[2] = g.dz;
  return r;
}

inline size_t n_vox(const muto_grid& g) {
  return (size_t)g.nx * g.ny * g.nz;
}

// RAII device buffer
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) count = 1;
    return hipMalloc((void**)&p, count * sizeof(T));
  }
  size_t bytes() const { return n * sizeof(T); }
};

// A copy ordered after the work already on the context's stream, waited
// for.  (The stream is non-blocking, so a plain hipMemcpy on the null
// stream would not wait for the stream's kernels.)
inline hipError_t copy_sync(muto_ctx* ctx, void* dst, const void* src, size_t bytes,
                            hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ctx->stream);
}

inline unsigned grid_for(size_t n, unsigned block) {
  size_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  if (g > 0x7fffffffull) g = 0x7fffffff;
  return (unsigned)g;
}

// Device address of this GPU's copy of the static scratch of
// muto_map_value_grad_noalloc, looked up once per device with the device
// current (map_op.hip); the status is kept for that entry point to report.
hipError_t noalloc_scratch_init(int device);

// Device-wide helpers (scan.hip)
hipError_t exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n,
                              hipStream_t s);
hipError_t exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out,
                                     size_t n, hipStream_t s);
hipError_t sort_pairs_u64_u32(uint64_t* keys, uint32_t* vals, size_t n,
                              int begin_bit, int end_bit, hipStream_t s);
hipError_t sort_pairs_u32_u32(uint32_t* keys, uint32_t* vals, size_t n,
                              int begin_bit, int end_bit, hipStream_t s);
hipError_t sum_f64(const double* in, double* out, size_t n, hipStream_t s);
// Workspace forms: no allocation and no synchronisation.  With temp ==
// nullptr they only set *temp_bytes.  The sorts leave the result in (keys,
// vals) or, when *in_alt, in (keys_alt, vals_alt).
hipError_t sort_pairs_u64_u32_ws(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                                 uint32_t* vals_alt, size_t n, int begin_bit, int end_bit,
                                 void* temp, size_t* temp_bytes, bool* in_alt, hipStream_t s);
hipError_t sort_pairs_u32_u32_ws(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                                 uint32_t* vals_alt, size_t n, int begin_bit, int end_bit,
                                 void* temp, size_t* temp_bytes, bool* in_alt, hipStream_t s);
hipError_t exclusive_scan_u32_to_u64_ws(const uint32_t* in, uint64_t* out, size_t n, void* temp,
                                        size_t* temp_bytes, hipStream_t s);

// Siddon over a batch of segments (siddon.hip)
hipError_t siddon_count(const GridD& g, const double* p1, const double* p2,
                        size_t n, uint64_t* counts, hipStream_t s);
// unpackable: null for the unpacked form; else the packed-staging form,
// which sets *unpackable when a recorded index does not fit 11/11/10 bits
hipError_t siddon_fill(const GridD& g, const double* p1, const double* p2,
                       size_t n, const uint64_t* row_ptr, uint32_t* vx,
                       uint32_t* vy, uint32_t* vz, double* len,
                       uint32_t* unpackable, hipStream_t s);

}  // namespace muto
