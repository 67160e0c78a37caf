// calib.hip -- the random row-gather ceiling of the device, for the roofline.
//
// The beam search's traffic is random rows: a visit reads one adjacency row
// and then one point row per evaluated neighbour, each a few 16-byte words
// at an unpredictable address.  Its bound is therefore not the streaming
// copy bandwidth but the rate at which the memory system serves independent
// cache lines.  k_calib_gather measures that rate: every 8-lane group reads
// read_bytes (<= 128) of a pseudo-random row of row_stride bytes in a buffer
// far larger than the caches, as the search's 8-lane gathers do.  Built as
// its own library (libpann_calib.so); nothing in libpann uses it.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__global__ void k_calib_gather(const char* buf, uint64_t nrows, uint32_t stride,
                               uint32_t read_bytes, uint64_t nreads, uint32_t* sink) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t sub = threadIdx.x & 7u;
  const uint64_t ngrp = ((uint64_t)gridDim.x * blockDim.x) >> 3;
  uint32_t acc = 0;
  for (uint64_t g = tid >> 3; g < nreads; g += ngrp) {
    uint64_t h = (g + 1) * 0x9E3779B97F4A7C15ull;  // splitmix64 finaliser
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    const uint64_t row = h % nrows;
    if (sub * 16u < read_bytes) {
      const uint4 v = *(const uint4*)(buf + row * stride + sub * 16u);
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
  }
  if (acc == 0x9E3779B1u) sink[0] = acc;  // keeps the loads; never true for a zeroed buffer
}

}  // namespace

extern "C" {

// Times one gather of nreads rows over buf (buf_bytes, 16-byte aligned,
// zero-filled by the caller) on `stream`; *ms = the kernel's duration.
// Returns 0, or a hipError_t value.
int pann_calib_gather(const void* buf, uint64_t buf_bytes, uint32_t row_stride,
                      uint32_t read_bytes, uint64_t nreads, uint32_t* sink, float* ms,
                      void* stream) {
  if (!buf || !sink || !ms || row_stride < 16 || row_stride % 16 || read_bytes == 0 ||
      read_bytes > 128 || read_bytes > row_stride || buf_bytes < 2ull * row_stride)
    return (int)hipErrorInvalidValue;
  const uint64_t nrows = buf_bytes / row_stride - 1;  // row * stride + row_stride <= buf_bytes
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_calib_gather, dim3(65536), dim3(256), 0, s, (const char*)buf, nrows,
                       row_stride, read_bytes, nreads, sink);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return (int)e;
}

}  // extern "C"
