// diag.hip — device diagnostics of libpcc.
//
// pcc_measure_copy_bandwidth / pcc_measure_read_bandwidth: the achievable
// HBM copy and read rates of this device, the practical ceilings the
// bandwidth-bound kernels are compared with beside the published peak.  The
// copy kernel, k_stream_copy, also serves as the copy of known size against
// which hardware traffic counters are calibrated.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace {

__global__ void __launch_bounds__(256) k_stream_copy(const double4* __restrict__ in,
                                                     double4* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = in[i];
}

// Sum of every double of `in` per thread (four loads in flight), one double
// out per thread so that no load is dead.
__global__ void __launch_bounds__(256) k_stream_read(const double4* __restrict__ in,
                                                     double* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  for (; i + 3 * stride < n; i += 4 * stride) {
    double4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = in[i + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
  }
  for (; i < n; i += stride) acc += (in[i].x + in[i].y) + (in[i].z + in[i].w);
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

constexpr int kReadGrid = 8192;

}  // namespace

extern "C" pcc_status pcc_measure_read_bandwidth(int device, int64_t bytes, int reps,
                                                 double* gbs) {
  if (!gbs || bytes < 1024 || reps < 1) return PCC_E_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return PCC_E_DEVICE;
  pcc::DeviceLock lock(pcc::device_mutex());
  const int64_t n = bytes / 32;  // double4 elements
  double4* a = nullptr;
  double* r = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  pcc_status st = PCC_OK;
  float ms = 0.f;
  if (hipMalloc(&a, 32 * n) != hipSuccess ||
      hipMalloc(&r, sizeof(double) * kReadGrid * 256) != hipSuccess) {
    st = PCC_E_OUT_OF_MEMORY;
    goto done;
  }
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipMemsetAsync(a, 0, 32 * n, s) != hipSuccess) {
    st = PCC_E_DEVICE;
    goto done;
  }
  k_stream_read<<<kReadGrid, 256, 0, s>>>(a, r, n);  // warm-up
  if (hipEventRecord(e0, s) != hipSuccess) {
    st = PCC_E_DEVICE;
    goto done;
  }
  for (int k = 0; k < reps; ++k) k_stream_read<<<kReadGrid, 256, 0, s>>>(a, r, n);
  if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess) {
    st = PCC_E_DEVICE;
    goto done;
  }
  *gbs = 32.0 * (double)n * reps / (ms * 1e-3) / 1e9;  // bytes read
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (r) (void)hipFree(r);
  return st;
}

extern "C" pcc_status pcc_measure_copy_bandwidth(int device, int64_t bytes, int reps,
                                                 double* gbs) {
  if (!gbs || bytes < 1024 || reps < 1) return PCC_E_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return PCC_E_DEVICE;
  pcc::DeviceLock lock(pcc::device_mutex());
  const int64_t n = bytes / 32;  // double4 elements
  double4* a = nullptr;
  double4* b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  pcc_status st = PCC_OK;
  float ms = 0.f;
  if (hipMalloc(&a, 32 * n) != hipSuccess || hipMalloc(&b, 32 * n) != hipSuccess) {
    st = PCC_E_OUT_OF_MEMORY;
    goto done;
  }
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipMemsetAsync(a, 0, 32 * n, s) != hipSuccess) {
    st = PCC_E_DEVICE;
    goto done;
  }
  k_stream_copy<<<8192, 256, 0, s>>>(a, b, n);  // warm-up
  if (hipEventRecord(e0, s) != hipSuccess) {
    st = PCC_E_DEVICE;
    goto done;
  }
  for (int r = 0; r < reps; ++r) k_stream_copy<<<8192, 256, 0, s>>>(a, b, n);
  if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess) {
    st = PCC_E_DEVICE;
    goto done;
  }
  *gbs = 2.0 * 32.0 * (double)n * reps / (ms * 1e-3) / 1e9;  // bytes read + written
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return st;
}
