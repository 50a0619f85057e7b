// Device-wide scans, sorts and sums used while building plans (set-up only,
// never inside an iteration).  hipCUB (header-only, rocPRIM backend).
#include <hipcub/hipcub.hpp>

#include "muto_internal.hpp"

namespace muto {

namespace {
struct Temp {
  void* p = nullptr;
  ~Temp() {
    if (p) (void)hipFree(p);
  }
};
struct U32ToU64 {
  __host__ __device__ uint64_t operator()(uint32_t v) const { return v; }
};
}  // namespace

hipError_t exclusive_scan_u64(const uint64_t* in, uint64_t* out, size_t n,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t bytes = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s);
  if (e != hipSuccess) return e;
  Temp t;
  if ((e = hipMalloc(&t.p, bytes)) != hipSuccess) return e;
  e = hipcub::DeviceScan::ExclusiveSum(t.p, bytes, in, out, n, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out,
                                     size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipcub::TransformInputIterator<uint64_t, U32ToU64, const uint32_t*> it(
      in, U32ToU64());
  size_t bytes = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, it, out, n, s);
  if (e != hipSuccess) return e;
  Temp t;
  if ((e = hipMalloc(&t.p, bytes)) != hipSuccess) return e;
  e = hipcub::DeviceScan::ExclusiveSum(t.p, bytes, it, out, n, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t sort_pairs_u64_u32(uint64_t* keys, uint32_t* vals, size_t n,
                              int begin_bit, int end_bit, hipStream_t s) {
  if (n <= 1) return hipSuccess;
  DBuf<uint64_t> k2;
  DBuf<uint32_t> v2;
  hipError_t e;
  if ((e = k2.alloc(n)) != hipSuccess) return e;
  if ((e = v2.alloc(n)) != hipSuccess) return e;
  hipcub::DoubleBuffer<uint64_t> kb(keys, k2.p);
  hipcub::DoubleBuffer<uint32_t> vb(vals, v2.p);
  size_t bytes = 0;
  e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kb, vb, (int)n,
                                         begin_bit, end_bit, s);
  if (e != hipSuccess) return e;
  Temp t;
  if ((e = hipMalloc(&t.p, bytes)) != hipSuccess) return e;
  e = hipcub::DeviceRadixSort::SortPairs(t.p, bytes, kb, vb, (int)n,
                                         begin_bit, end_bit, s);
  if (e != hipSuccess) return e;
  if (kb.Current() != keys) {
    if ((e = hipMemcpyAsync(keys, kb.Current(), n * sizeof(uint64_t),
                            hipMemcpyDeviceToDevice, s)) != hipSuccess)
      return e;
  }
  if (vb.Current() != vals) {
    if ((e = hipMemcpyAsync(vals, vb.Current(), n * sizeof(uint32_t),
                            hipMemcpyDeviceToDevice, s)) != hipSuccess)
      return e;
  }
  return hipStreamSynchronize(s);
}

hipError_t sort_pairs_u32_u32(uint32_t* keys, uint32_t* vals, size_t n,
                              int begin_bit, int end_bit, hipStream_t s) {
  if (n <= 1) return hipSuccess;
  DBuf<uint32_t> k2, v2;
  hipError_t e;
  if ((e = k2.alloc(n)) != hipSuccess) return e;
  if ((e = v2.alloc(n)) != hipSuccess) return e;
  hipcub::DoubleBuffer<uint32_t> kb(keys, k2.p);
  hipcub::DoubleBuffer<uint32_t> vb(vals, v2.p);
  size_t bytes = 0;
  e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kb, vb, (int)n,
                                         begin_bit, end_bit, s);
  if (e != hipSuccess) return e;
  Temp t;
  if ((e = hipMalloc(&t.p, bytes)) != hipSuccess) return e;
  e = hipcub::DeviceRadixSort::SortPairs(t.p, bytes, kb, vb, (int)n,
                                         begin_bit, end_bit, s);
  if (e != hipSuccess) return e;
  if (kb.Current() != keys) {
    if ((e = hipMemcpyAsync(keys, kb.Current(), n * sizeof(uint32_t),
                            hipMemcpyDeviceToDevice, s)) != hipSuccess)
      return e;
  }
  if (vb.Current() != vals) {
    if ((e = hipMemcpyAsync(vals, vb.Current(), n * sizeof(uint32_t),
                            hipMemcpyDeviceToDevice, s)) != hipSuccess)
      return e;
  }
  return hipStreamSynchronize(s);
}

hipError_t sum_f64(const double* in, double* out, size_t n, hipStream_t s) {
  size_t bytes = 0;
  hipError_t e = hipcub::DeviceReduce::Sum(nullptr, bytes, in, out, (int)n, s);
  if (e != hipSuccess) return e;
  Temp t;
  if ((e = hipMalloc(&t.p, bytes)) != hipSuccess) return e;
  e = hipcub::DeviceReduce::Sum(t.p, bytes, in, out, (int)n, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

template <class K>
static hipError_t sort_pairs_ws(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, size_t n,
                                int begin_bit, int end_bit, void* temp, size_t* temp_bytes,
                                bool* in_alt, hipStream_t s) {
  if (in_alt) *in_alt = false;
  if (n <= 1) {
    *temp_bytes = 16;
    return hipSuccess;
  }
  hipcub::DoubleBuffer<K> kb(keys, keys_alt);
  hipcub::DoubleBuffer<uint32_t> vb(vals, vals_alt);
  if (!temp)
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *temp_bytes, kb, vb, (int)n, begin_bit,
                                              end_bit, s);
  const hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, kb, vb, (int)n,
                                                          begin_bit, end_bit, s);
  if (in_alt) *in_alt = kb.Current() != keys;
  return e;
}

hipError_t sort_pairs_u64_u32_ws(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                                 uint32_t* vals_alt, size_t n, int begin_bit, int end_bit,
                                 void* temp, size_t* temp_bytes, bool* in_alt, hipStream_t s) {
  return sort_pairs_ws(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, temp, temp_bytes,
                       in_alt, s);
}

hipError_t sort_pairs_u32_u32_ws(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                                 uint32_t* vals_alt, size_t n, int begin_bit, int end_bit,
                                 void* temp, size_t* temp_bytes, bool* in_alt, hipStream_t s) {
  return sort_pairs_ws(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, temp, temp_bytes,
                       in_alt, s);
}

hipError_t exclusive_scan_u32_to_u64_ws(const uint32_t* in, uint64_t* out, size_t n, void* temp,
                                        size_t* temp_bytes, hipStream_t s) {
  if (n == 0) {
    *temp_bytes = 16;
    return hipSuccess;
  }
  hipcub::TransformInputIterator<uint64_t, U32ToU64, const uint32_t*> it(in, U32ToU64());
  return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, it, out, n, s);
}

}  // namespace muto
