// hbm_copy.hip -- reference copy kernel for the HBM byte counters.
//
// Copies N MiB (default 256) from one device buffer to another with 16-byte
// vector loads and stores, once, after an untimed fill.  The kernel reads and
// writes exactly N MiB, so the FETCH_SIZE / WRITE_SIZE counters that a
// profiler reports for its dispatch give the scale from counter units to
// bytes moved for the other kernels of this library.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void fill_words(int4* p, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    p[e] = make_int4((int)e, (int)(e >> 32), 1, 2);
}

__global__ void copy_words(const int4* __restrict__ src, int4* __restrict__ dst, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

int main(int argc, char** argv) {
  const long mib = argc > 1 ? atol(argv[1]) : 256;
  if (mib < 1 || mib > 65536) {
    fprintf(stderr, "hbm_copy: size must be 1..65536 MiB\n");
    return 2;
  }
  const size_t bytes = (size_t)mib << 20, n = bytes / sizeof(int4);
  int4 *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) {
    fprintf(stderr, "hbm_copy: allocation failed\n");
    return 1;
  }
  hipLaunchKernelGGL(fill_words, dim3(4096), dim3(256), 0, 0, src, n);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(copy_words, dim3(8192), dim3(256), 0, 0, src, dst, n);
  (void)hipEventRecord(e1, 0);
  if (hipEventSynchronize(e1) != hipSuccess) {
    fprintf(stderr, "hbm_copy: kernel failed\n");
    return 1;
  }
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  printf("hbm_copy: %zu bytes read, %zu bytes written, %.3f ms, %.1f GB/s\n", bytes, bytes, ms, 2.0 * bytes / (ms * 1e6));
  (void)hipFree(src);
  (void)hipFree(dst);
  return 0;
}
