// hbm_copy.hip -- calibration of the HBM traffic counters: one plain copy of
// 1 GiB (FP64, one element per thread), then a 1 GiB read-only pass.  Under
// `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` the per-dispatch counter
// values of these kernels give the scale from counter units to bytes.
// Then the measured copy peak: the best of the plain copy (warm) and of a
// 1 GiB copy with 16-byte accesses, four per thread in flight, over grids
// of 4 to 32 blocks per CU, printed as `copy_peak_GBs <value>` (bytes read
// plus bytes written per second).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e = (x);                                                         \
    if (e != hipSuccess) {                                                      \
      fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e), __LINE__); \
      return 1;                                                                 \
    }                                                                           \
  } while (0)

__global__ void copy_f64(const double* __restrict__ a, double* __restrict__ b, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void read_f64(const double* __restrict__ a, double* __restrict__ sink, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n && a[i] == 12345.678) sink[0] = a[i];
}

typedef double d2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) copy_wide(const d2* __restrict__ a, d2* __restrict__ b, size_t n2) {
  // each block copies runs of 4 x 256 consecutive 16-byte elements, one per
  // thread and pass, every pass coalesced; blocks stride over the array
  for (size_t base = (size_t)blockIdx.x * 1024; base < n2; base += (size_t)gridDim.x * 1024) {
    d2 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t i = base + k * 256 + threadIdx.x;
      if (i < n2) v[k] = a[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t i = base + k * 256 + threadIdx.x;
      if (i < n2) b[i] = v[k];
    }
  }
}

int main() {
  const size_t bytes = (size_t)1 << 30;
  const size_t n = bytes / sizeof(double);
  double *a = nullptr, *b = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMemset(a, 0, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const unsigned blocks = (unsigned)((n + 255) / 256);
  for (int rep = 0; rep < 4; ++rep) {
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(copy_f64, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("copy 1 GiB: %.3f ms, %.1f GB/s (read + write)\n", ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
  }
  for (int rep = 0; rep < 2; ++rep) {
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(read_f64, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("read 1 GiB: %.3f ms, %.1f GB/s\n", ms, 1.0 * bytes / (ms * 1e-3) / 1e9);
  }
  int dev = 0, cus = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  float best = 1e30f;
  for (int rep = 0; rep < 3; ++rep) {   // the plain copy again, warm
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(copy_f64, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) best = ms;
  }
  for (int occ : {4, 8, 16, 32}) {
    const unsigned g = (unsigned)(cus > 0 ? cus : 256) * occ;
    for (int rep = 0; rep < 5; ++rep) {
      CK(hipEventRecord(e0));
      hipLaunchKernelGGL(copy_wide, dim3(g), dim3(256), 0, 0, (const d2*)a, (d2*)b, n / 2);
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0 && ms < best) best = ms;
    }
  }
  printf("copy_peak_GBs %.1f\n", 2.0 * bytes / (best * 1e-3) / 1e9);
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
