// Peak probes for roofline reports: the FP64 FMA rate and the HBM copy
// bandwidth this GPU actually reaches, measured with two plain kernels.
#include <hip/hip_runtime.h>

#include "../../include/meamz.h"

namespace {

// Eight independent FMA chains per lane, 2 FLOP each per iteration.  The
// store is taken only for a value the chains never reach, so the loop stays
// and nothing is written.
__global__ void __launch_bounds__(256) k_fma_probe(double* sink, double seed, int iters) {
  const double b = 0.9999999, c = 1e-9;
  double a[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) a[q] = seed + 1e-12 * (double)(threadIdx.x + 37 * q);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = fma(a[q], b, c);
  }
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) s += a[q];
  if (s == -1.0) sink[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

typedef double dvec2 __attribute__((ext_vector_type(2)));

// Stream copy, grid-stride, four independent 16-byte non-temporal loads per
// lane in flight before their stores.
__global__ void __launch_bounds__(256) k_copy_probe(const dvec2* __restrict__ src, dvec2* __restrict__ dst,
                                                     size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; q + 3 * stride < n; q += 4 * stride) {
    const dvec2 a = __builtin_nontemporal_load(src + q);
    const dvec2 b = __builtin_nontemporal_load(src + q + stride);
    const dvec2 c = __builtin_nontemporal_load(src + q + 2 * stride);
    const dvec2 d = __builtin_nontemporal_load(src + q + 3 * stride);
    __builtin_nontemporal_store(a, dst + q);
    __builtin_nontemporal_store(b, dst + q + stride);
    __builtin_nontemporal_store(c, dst + q + 2 * stride);
    __builtin_nontemporal_store(d, dst + q + 3 * stride);
  }
  for (; q < n; q += stride) dst[q] = src[q];
}

}  // namespace

extern "C" int mz_probe_peaks(double* fp64_tflops, double* copy_gbs) {
  if (!fp64_tflops || !copy_gbs) return MZ_ERR_ARG;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return MZ_ERR_HIP;
  hipStream_t st;
  if (hipStreamCreate(&st) != hipSuccess) return MZ_ERR_HIP;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const size_t bytes = (size_t)1 << 30;  // 1 GiB each way
  const size_t n = bytes / sizeof(dvec2);
  dvec2 *src = nullptr, *dst = nullptr;
  double* sink = nullptr;
  const int blocks = prop.multiProcessorCount * 32, threads = 256, iters = 8192;
  int rc = MZ_OK;
  if (hipMalloc((void**)&src, bytes) != hipSuccess || hipMalloc((void**)&dst, bytes) != hipSuccess ||
      hipMalloc((void**)&sink, sizeof(double) * (size_t)blocks * threads) != hipSuccess) {
    rc = MZ_ERR_NOMEM;
  } else {
    (void)hipMemsetAsync(src, 0, bytes, st);
    float best_f = 1e30f, best_c = 1e30f;
    for (int rep = 0; rep < 4; ++rep) {  // the first of each is a warm-up
      float ms = 0.f;
      (void)hipEventRecord(e0, st);
      hipLaunchKernelGGL(k_fma_probe, dim3(blocks), dim3(threads), 0, st, sink, 1.0, iters);
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0 && ms < best_f) best_f = ms;
      for (int wg = 4; wg <= 32; wg *= 2) {  // workgroups per CU
        (void)hipEventRecord(e0, st);
        hipLaunchKernelGGL(k_copy_probe, dim3(prop.multiProcessorCount * wg), dim3(256), 0, st, src, dst, n);
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < best_c) best_c = ms;
      }
    }
    if (hipGetLastError() != hipSuccess) rc = MZ_ERR_HIP;
    const double flops = 2.0 * 8.0 * iters * (double)blocks * threads;
    *fp64_tflops = flops / (best_f * 1e-3) / 1e12;
    *copy_gbs = 2.0 * (double)bytes / (best_c * 1e-3) / 1e9;
  }
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (sink) (void)hipFree(sink);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  return rc;
}
