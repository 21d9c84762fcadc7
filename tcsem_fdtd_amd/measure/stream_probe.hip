// stream_probe.hip -- HBM streaming probes: the bandwidth ceiling that the volume march is measured
// against, and the calibration kernel of known traffic for the FETCH_SIZE / WRITE_SIZE counters.
//
//   stream_probe copy MIB REPS          one plain copy (16-byte loads and stores, grid-stride, one
//                                       load in flight per lane): the counter calibration kernel
//   stream_probe sweep [CELLS] [REPS]   every probe below over CELLS elements per stream (default
//                                       67108864, the 512 x 512 x 256 model's cell count)
//
// Probes (read+write bytes over the launch time from HIP events; best and mean of REPS launches):
//   copy4_u1       b = a, float4 per lane, one load in flight per lane (the plain copy)
//   copy4_u4_nt    four 16-byte loads in flight per lane before the stores, non-temporal stores
//   copy1_u4_nt    the same with 4-byte lanes
//   read6          6 arrays read and reduced (no stores): the read-only ceiling
//   mix11_fW_uU    the march's stream mix: 6 arrays read, 5 written, W floats per lane per stream
//                  (W = 1, 2, 4), U elements in flight per lane, non-temporal stores
//   mix5_fW_uU     3 arrays read, 2 written
// Streams sit in one allocation; "_stag" probes shift stream q by q x (4 KiB + 256 B) beyond
// 2 MiB-aligned bases.
// Each probe is launched with 8 workgroups of 256 threads per CU (32 waves per CU).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#define CHK(x)                                                                    \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
      std::exit(1);                                                               \
    }                                                                             \
  } while (0)

__global__ void __launch_bounds__(256) k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    b[t] = a[t];
}

typedef float nf4 __attribute__((ext_vector_type(4)));  // native vectors for the NT builtin
typedef float nf2 __attribute__((ext_vector_type(2)));
template <typename V, bool NT>
__device__ __forceinline__ void put(V *p, V v) {
  if constexpr (!NT) {
    *p = v;
  } else if constexpr (sizeof(V) == 16) {
    nf4 w;
    __builtin_memcpy(&w, &v, 16);
    __builtin_nontemporal_store(w, (nf4 *)p);
  } else if constexpr (sizeof(V) == 8) {
    nf2 w;
    __builtin_memcpy(&w, &v, 8);
    __builtin_nontemporal_store(w, (nf2 *)p);
  } else {
    __builtin_nontemporal_store(v, p);
  }
}

// U loads in flight per lane: the U loads of one round are issued before its U stores.
template <typename V, int U, bool NT>
__global__ void __launch_bounds__(256) k_copy_u(const V *__restrict__ a, V *__restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t + (U - 1) * stride < n; t += U * stride) {
    V v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = a[t + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) put<V, NT>(&b[t + u * stride], v[u]);
  }
  for (; t < n; t += stride) put<V, NT>(&b[t], a[t]);
}

template <typename V>
__device__ __forceinline__ V axpy(V x, V y, float s);
template <>
__device__ __forceinline__ float axpy(float x, float y, float s) {
  return x + s * y;
}
template <>
__device__ __forceinline__ float2 axpy(float2 x, float2 y, float s) {
  return make_float2(x.x + s * y.x, x.y + s * y.y);
}
template <>
__device__ __forceinline__ float4 axpy(float4 x, float4 y, float s) {
  return make_float4(x.x + s * y.x, x.y + s * y.y, x.z + s * y.z, x.w + s * y.w);
}

struct Mix {
  const void *in[6];
  void *out[5];
};

// The march's mix: reads sigma, Ex, Ey, Ez, Bx, By; writes Ex, Ey, Ez, Bx, By (each output from
// its own input and one other stream, so that no load is dead).  NI inputs, NI - 1 outputs.
template <typename V, int U, int NI = 6>
__global__ void __launch_bounds__(256) k_mix(const Mix m, size_t n, float s) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const V *in[NI];
  V *out[NI - 1];
#pragma unroll
  for (int q = 0; q < NI; ++q) in[q] = (const V *)m.in[q];
#pragma unroll
  for (int q = 0; q < NI - 1; ++q) out[q] = (V *)m.out[q];
  for (; t + (U - 1) * stride < n; t += U * stride) {
    V v[U][NI];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < NI; ++q) v[u][q] = in[q][t + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < NI - 1; ++q) put<V, true>(&out[q][t + u * stride], axpy(v[u][q + 1], v[u][0], s));
  }
  for (; t < n; t += stride)
#pragma unroll
    for (int q = 0; q < NI - 1; ++q) put<V, true>(&out[q][t], axpy(in[q + 1][t], in[0][t], s));
}

__global__ void __launch_bounds__(256) k_read6(const Mix m, size_t n, float *sink) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const float4 *in[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) in[q] = (const float4 *)m.in[q];
  float acc = 0.f;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const float4 v = in[q][t];
      acc += v.x + v.y + v.z + v.w;
    }
  if (acc == 12345.678f) sink[0] = acc;  // keeps the loads live; never true for the fill values
}

__global__ void k_fill(float *a, size_t n, float v) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    a[t] = v + (float)(t & 1023) * 1e-3f;
}

struct Timer {
  hipEvent_t e0, e1;
  Timer() {
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
  }
  ~Timer() {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
};

template <typename F>
void timed(const char *name, double bytes, int reps, F launch) {
  Timer t;
  launch();  // first launch: code object load
  CHK(hipDeviceSynchronize());
  double best = 0, sum = 0;
  for (int r = 0; r < reps; ++r) {
    CHK(hipEventRecord(t.e0, 0));
    launch();
    CHK(hipEventRecord(t.e1, 0));
    CHK(hipEventSynchronize(t.e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, t.e0, t.e1));
    const double gbs = bytes / (ms * 1e-3) / 1e9;
    sum += gbs;
    if (gbs > best) best = gbs;
  }
  std::printf("{\"probe\": \"%s\", \"bytes\": %.0f, \"reps\": %d, \"best_GBs\": %.1f, \"mean_GBs\": %.1f}\n", name,
              bytes, reps, best, sum / reps);
  std::fflush(stdout);
}

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "sweep";
  int dev = 0, cu = 0;
  CHK(hipGetDevice(&dev));
  CHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
  const dim3 grd(cu * 8), blk(256);
  if (mode == "copy") {
    const size_t mib = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 2048;
    const int reps = argc > 3 ? std::atoi(argv[3]) : 20;
    const size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, bytes));
    CHK(hipMalloc(&b, bytes));
    hipLaunchKernelGGL(k_fill, grd, blk, 0, 0, a, bytes / 4, 1.f);
    timed("copy4_u1", 2.0 * bytes, reps,
          [&] { hipLaunchKernelGGL(k_copy, grd, blk, 0, 0, (const float4 *)a, (float4 *)b, n); });
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
  }
  if (mode != "sweep") {
    std::fprintf(stderr, "usage: stream_probe copy MIB REPS | sweep [CELLS] [REPS]\n");
    return 2;
  }
  size_t cells = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : ((size_t)1 << 26);
  cells &= ~(size_t)255;  // whole float4 vectors per stream
  const int reps = argc > 3 ? std::atoi(argv[3]) : 10;
  // 11 streams in one allocation: "aligned" bases at multiples of 2 MiB, "stag" bases shifted by
  // a further q x (4 KiB + 256 B) so that equal offsets of different streams fall on different
  // memory channels
  const size_t span = ((cells * sizeof(float) + (2u << 20) - 1) / (2u << 20)) * (2u << 20);
  const size_t stag = 4096 + 256;
  char *buf = nullptr;
  CHK(hipMalloc(&buf, 11 * span + 11 * stag + 4096));
  hipLaunchKernelGGL(k_fill, grd, blk, 0, 0, (float *)buf, (11 * span + 11 * stag) / 4, 1.f);
  float *sink = nullptr;
  CHK(hipMalloc(&sink, 64));
  CHK(hipDeviceSynchronize());
  std::printf("{\"cu\": %d, \"cells\": %zu, \"grid\": %u, \"block\": %u, \"stagger_bytes\": %zu}\n", cu, cells,
              grd.x, blk.x, stag);
  auto mix_of = [&](bool st) {
    Mix m;
    for (int q = 0; q < 6; ++q) m.in[q] = buf + q * span + (st ? q * stag : 0);
    for (int q = 0; q < 5; ++q) m.out[q] = buf + (6 + q) * span + (st ? (6 + q) * stag : 0);
    return m;
  };
  const Mix ma = mix_of(false), ms = mix_of(true);
  const double b4 = (double)cells * sizeof(float);
  const size_t n4 = cells / 4, n2 = cells / 2;
  timed("copy4_u1", 2 * b4, reps, [&] { hipLaunchKernelGGL(k_copy, grd, blk, 0, 0, (const float4 *)ma.in[0], (float4 *)ma.out[0], n4); });
  timed("copy4_u4_nt", 2 * b4, reps, [&] {
    hipLaunchKernelGGL((k_copy_u<float4, 4, true>), grd, blk, 0, 0, (const float4 *)ma.in[0], (float4 *)ma.out[0], n4);
  });
  timed("copy1_u4_nt", 2 * b4, reps, [&] {
    hipLaunchKernelGGL((k_copy_u<float, 4, true>), grd, blk, 0, 0, (const float *)ma.in[0], (float *)ma.out[0], cells);
  });
  timed("read6", 6 * b4, reps, [&] { hipLaunchKernelGGL(k_read6, grd, blk, 0, 0, ma, n4, sink); });
  timed("mix11_f1_u2", 11 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float, 2>), grd, blk, 0, 0, ma, cells, 0.5f); });
  timed("mix11_f1_u2_stag", 11 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float, 2>), grd, blk, 0, 0, ms, cells, 0.5f); });
  timed("mix11_f1_u4_stag", 11 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float, 4>), grd, blk, 0, 0, ms, cells, 0.5f); });
  timed("mix11_f2_u2", 11 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float2, 2>), grd, blk, 0, 0, ma, n2, 0.5f); });
  timed("mix11_f2_u2_stag", 11 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float2, 2>), grd, blk, 0, 0, ms, n2, 0.5f); });
  timed("mix11_f4_u2", 11 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float4, 2>), grd, blk, 0, 0, ma, n4, 0.5f); });
  timed("mix11_f4_u2_stag", 11 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float4, 2>), grd, blk, 0, 0, ms, n4, 0.5f); });
  timed("mix5_f1_u2_stag", 5 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float, 2, 3>), grd, blk, 0, 0, ms, cells, 0.5f); });
  timed("mix5_f4_u2_stag", 5 * b4, reps, [&] { hipLaunchKernelGGL((k_mix<float4, 2, 3>), grd, blk, 0, 0, ms, n4, 0.5f); });
  CHK(hipFree(buf));
  CHK(hipFree(sink));
  return 0;
}
