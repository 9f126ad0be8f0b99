// Measured ceilings of the device the path runs on: a plain device copy
// (HBM read + write) and FP64 / FP32 fused multiply-add throughput, timed
// with HIP events.  They put the kernels' roofline fractions beside the
// published MI355X peaks (HBM 8 TB/s, FP64 vector 78.6 TFLOP/s, FP32 vector
// 157.3 TFLOP/s, the latter counting packed FP32 instructions, which this
// library's build does not select).

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace {

__global__ void __launch_bounds__(256) copy_kernel(size_t n4, const double4 *__restrict__ a, double4 *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// kChains independent FMA chains per thread, kIters rounds each
constexpr int kChains = 8, kIters = 4096;
template <class T>
__global__ void __launch_bounds__(256) fma_kernel(T b, T c, T *__restrict__ out) {
  T a[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) a[k] = T(threadIdx.x + k);
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = fma(a[k], b, c);
  T s = T(0);
#pragma unroll
  for (int k = 0; k < kChains; ++k) s += a[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <class T>
int time_fma(hipStream_t s, hipEvent_t e0, hipEvent_t e1, double *tflops) {
  const unsigned blocks = 256 * 8 * 4;  // 4 workgroups of 256 per SIMD-quad on 256 CUs, 8 rounds
  crimson::DBuf<T> out;
  CR_TRY(out.alloc((size_t)blocks * 256));
  fma_kernel<T><<<blocks, 256, 0, s>>>(T(0.999), T(1e-3), out.p);  // warm-up
  CR_HIP(hipEventRecord(e0, s));
  const int reps = 5;
  for (int r = 0; r < reps; ++r) fma_kernel<T><<<blocks, 256, 0, s>>>(T(0.999), T(1e-3), out.p);
  CR_HIP(hipEventRecord(e1, s));
  CR_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CR_HIP(hipEventElapsedTime(&ms, e0, e1));
  const double flops = 2.0 * kChains * kIters * (double)blocks * 256 * reps;
  *tflops = flops / (ms * 1e-3) / 1e12;
  return CRIMSON_OK;
}

}  // namespace

extern "C" int crimson_measure_peaks(size_t copy_bytes, double *copy_gbs, double *fp64_tflops, double *fp32_tflops) {
  if (!copy_gbs || !fp64_tflops || !fp32_tflops || copy_bytes < (1u << 20)) {
    crimson::set_error("crimson_measure_peaks: invalid arguments");
    return CRIMSON_EARG;
  }
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CR_HIP(hipStreamCreate(&s));
  int rc = CRIMSON_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    crimson::set_error("hipEventCreate failed");
    rc = CRIMSON_EHIP;
  }
  if (rc == CRIMSON_OK) {
    const size_t n4 = copy_bytes / sizeof(double4);
    crimson::DBuf<double4> a, b;
    if ((rc = a.alloc(n4)) == CRIMSON_OK && (rc = b.alloc(n4)) == CRIMSON_OK) {
      const unsigned grid = 256 * 32;
      hipMemsetAsync(a.p, 0, n4 * sizeof(double4), s);
      copy_kernel<<<grid, 256, 0, s>>>(n4, a.p, b.p);
      hipEventRecord(e0, s);
      const int reps = 10;
      for (int r = 0; r < reps; ++r) copy_kernel<<<grid, 256, 0, s>>>(n4, a.p, b.p);
      hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) {
        crimson::set_error("copy probe failed");
        rc = CRIMSON_EHIP;
      } else {
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        *copy_gbs = 2.0 * n4 * sizeof(double4) * reps / (ms * 1e-3) / 1e9;
      }
    }
    if (rc == CRIMSON_OK) rc = time_fma<double>(s, e0, e1, fp64_tflops);
    if (rc == CRIMSON_OK) rc = time_fma<float>(s, e0, e1, fp32_tflops);
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(s);
  hipStreamDestroy(s);
  return rc;
}
