// Counter calibration: a plain device-to-device copy of known size (float4, grid-stride, coalesced).
// A profiler's FETCH_SIZE / WRITE_SIZE for this kernel against its known bytes gives the factor that turns
// those counters into HBM bytes for the step kernels.  Not part of the step library.
#include <hip/hip_runtime.h>
#include <stddef.h>

__global__ void __launch_bounds__(256) k_calib_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_calib_fill(float4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = make_float4(1.f, 2.f, 3.f, (float)(i & 1023));
}

// Copies `nbytes` (a multiple of 16) `reps` times between two fresh device buffers; 0 on success.
extern "C" int steel_calib_copy(size_t nbytes, int reps) {
  if (nbytes == 0 || nbytes % 16) return -3;
  float4 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, nbytes) != hipSuccess) return -4;
  if (hipMalloc(&b, nbytes) != hipSuccess) { (void)hipFree(a); return -4; }
  size_t n = nbytes / 16;
  hipLaunchKernelGGL(k_calib_fill, dim3(4096), dim3(256), 0, 0, a, n);
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_calib_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
  int rc = hipDeviceSynchronize() == hipSuccess ? 0 : -1;
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
