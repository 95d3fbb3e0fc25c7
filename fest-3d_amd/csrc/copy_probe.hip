// copy_probe.hip -- a plain device copy of known size, the calibration run
// for the HBM byte counters (FETCH_SIZE, WRITE_SIZE): the counters of this
// kernel against its 512 MiB read + 512 MiB write give the scale factor
// applied to the counters of the engine's kernels.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void __launch_bounds__(256) k_copy_probe(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                    long n) {
  long stride = (long)gridDim.x * blockDim.x;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) dst[t] = src[t];
}

int main() {
  const long n = 32L << 20;  // double2 elements: 512 MiB per buffer
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess) {
    fprintf(stderr, "copy_probe: hipMalloc failed\n");
    return 1;
  }
  if (hipMemset(a, 0, n * sizeof(double2)) != hipSuccess) return 1;
  for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, 0, a, b, n);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipGetLastError();
  printf("copy_probe: %ld bytes read and %ld bytes written per launch, 5 launches: %s\n", n * (long)sizeof(double2),
         n * (long)sizeof(double2), hipGetErrorString(e));
  (void)hipFree(a);
  (void)hipFree(b);
  return e == hipSuccess ? 0 : 1;
}
