// hbm_copy.hip -- stream-copy probe of the device's HBM bandwidth.
//
// Copies a buffer of `mib` MiB into another `reps` times with 16-byte vector loads and stores
// (grid-stride, 8 waves per CU), timing each launch with HIP events.  It reports read+write GB/s
// and is the calibration kernel of known traffic for the FETCH_SIZE / WRITE_SIZE counters.
//   usage: hbm_copy [mib=2048] [reps=20]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    b[t] = a[t];
}

__global__ void k_fill(float4 *a, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    a[t] = make_float4(1.f, 2.f, 3.f, (float)(t & 1023));
}

#define CHK(x)                                                                    \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

int main(int argc, char **argv) {
  const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 2048;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  const size_t bytes = mib << 20, n = bytes / sizeof(float4);
  int dev = 0, cu = 0;
  CHK(hipGetDevice(&dev));
  CHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  float4 *a = nullptr, *b = nullptr;
  CHK(hipMalloc(&a, bytes));
  CHK(hipMalloc(&b, bytes));
  const dim3 grd(cu * 8), blk(256);
  hipLaunchKernelGGL(k_fill, grd, blk, 0, 0, a, n);
  hipLaunchKernelGGL(k_copy, grd, blk, 0, 0, a, b, n);
  CHK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  CHK(hipEventCreate(&e0));
  CHK(hipEventCreate(&e1));
  double best = 0, sum = 0;
  for (int r = 0; r < reps; ++r) {
    CHK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_copy, grd, blk, 0, 0, a, b, n);
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double gbs = 2.0 * bytes / (ms * 1e-3) / 1e9;
    sum += gbs;
    if (gbs > best) best = gbs;
  }
  std::printf("{\"copy_bytes\": %zu, \"reps\": %d, \"cu\": %d, \"best_GBs\": %.1f, \"mean_GBs\": %.1f}\n", bytes,
              reps, cu, best, sum / reps);
  CHK(hipFree(a));
  CHK(hipFree(b));
  return 0;
}
