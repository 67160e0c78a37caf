// quantize.hip -- reduced-precision copies of an FP32 point store (K8).
//
// BF16: round-to-nearest-even of each f32 (NaN inputs are not expected).
// FP8:  OCP e4m3fn with a per-dimension scale s_j = max_i |x_ij| / 448
//       (448 = largest finite e4m3), code = RNE(x_ij / s_j) saturated to
//       +-448; subnormals (|v| < 2^-6) are kept (step 2^-9).
// FP8 residual: the same codes of x_i - c_{a(i)}, c the nearest sampled
//       point, the scales taken over the residuals; the centroid id a(i) sits
//       in the row's last 4 bytes.
// The reference's analog is its scalar quantization of points
// (euclidian_point.h:188-241, mips_point.h:425-501), which uses one global
// scale; this store scales per dimension instead.
#include <algorithm>

#include "kernels.h"
#include "device_common.h"
#include "dist.h"

namespace pann {

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint8_t f32_to_e4m3_rne(float v) {
  uint32_t sgn = (__float_as_uint(v) >> 24) & 0x80u;
  float a = fabsf(v);
  uint32_t code;
  if (!(a < 448.0f)) {
    code = 0x7Eu;  // saturate (also catches +inf)
  } else if (a < 0.015625f) {  // below 2^-6: subnormal grid m * 2^-9, m in [0, 8]
    code = (uint32_t)rintf(a * 512.0f);  // exact scaling, RNE; m == 8 is the first normal
  } else {
    uint32_t b = __float_as_uint(a);
    b += 0x7FFFFu + ((b >> 20) & 1u);      // RNE to 3 mantissa bits
    code = (b >> 20) - ((127u - 7u) << 3); // re-bias the exponent 127 -> 7
    if (code > 0x7Eu) code = 0x7Eu;
  }
  return (uint8_t)(code | sgn);
}

// The value encoded for row r, column j: x_rj, or x_rj - c_{assign[r], j}
// for the residual store.
__device__ __forceinline__ float store_value(const float* X, uint64_t stride_f, const float* cent,
                                             uint64_t cstride_f, const uint32_t* assign,
                                             uint64_t r, uint64_t j) {
  float v = X[r * stride_f + j];
  if (cent) v -= cent[(uint64_t)assign[r] * cstride_f + j];
  return v;
}

__global__ void k_absmax_cols(const float* __restrict__ X, uint64_t stride_f, uint32_t n,
                              uint32_t dim, uint32_t* __restrict__ out_bits,
                              const float* __restrict__ cent, uint64_t cstride_f,
                              const uint32_t* __restrict__ assign) {
  // each block: a slab of rows, one thread per column (looping)
  uint32_t r0 = blockIdx.x * 256u;
  uint32_t r1 = r0 + 256u < n ? r0 + 256u : n;
  for (uint32_t j = threadIdx.x; j < dim; j += blockDim.x) {
    float m = 0.0f;
    for (uint32_t r = r0; r < r1; r++)
      m = fmaxf(m, fabsf(store_value(X, stride_f, cent, cstride_f, assign, r, j)));
    atomicMax(&out_bits[j], __float_as_uint(m));  // non-negative floats order as uints
  }
}

__global__ void k_scales_from_absmax(const uint32_t* __restrict__ bits, uint32_t dim,
                                     float* __restrict__ scales) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < dim) {
    float m = __uint_as_float(bits[j]);
    scales[j] = m > 0.0f ? m / 448.0f : 1.0f;
  }
}

__global__ void k_encode_bf16(const float* __restrict__ X, uint64_t stride_f, uint32_t n,
                              uint32_t dim, uint16_t* __restrict__ Y, uint64_t stride_h) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t total = (uint64_t)n * dim;
  if (i >= total) return;
  uint64_t r = i / dim, j = i % dim;
  Y[r * stride_h + j] = f32_to_bf16_rne(X[r * stride_f + j]);
}

__global__ void k_encode_fp8(const float* __restrict__ X, uint64_t stride_f, uint32_t n,
                             uint32_t dim, const float* __restrict__ scales,
                             uint8_t* __restrict__ Y, uint64_t stride_b,
                             const float* __restrict__ cent, uint64_t cstride_f,
                             const uint32_t* __restrict__ assign) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t total = (uint64_t)n * dim;
  if (i >= total) return;
  uint64_t r = i / dim, j = i % dim;
  Y[r * stride_b + j] =
      f32_to_e4m3_rne(store_value(X, stride_f, cent, cstride_f, assign, r, j) / scales[j]);
}

// The residual store's centroid id in the row's last 4 bytes, one 6-bit digit
// per byte (least significant first): every byte is a finite e4m3 code
// (<= 0x3F), so the search may decode it with the row; the query is zero
// there.
__global__ void k_put_centroid_ids(uint8_t* __restrict__ Y, uint64_t stride_b, uint32_t n,
                                   const uint32_t* __restrict__ assign) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint32_t c = assign[r];
  uint32_t w = (c & 63u) | ((c >> 6) & 63u) << 8 | ((c >> 12) & 63u) << 16 | ((c >> 18) & 63u) << 24;
  *(uint32_t*)(Y + (uint64_t)r * stride_b + stride_b - 4) = w;
}

__global__ void k_zero_u32_quant(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Decodes all 256 e4m3 codes with the search kernels' decoder (for tests).
__global__ void k_fp8_decode_all(float* out) {
  uint32_t i = threadIdx.x;
  if (i < 256u) out[i] = fp8e4m3_to_f32(i);
}

// Encodes n floats with the store encoder (for tests).
__global__ void k_fp8_encode_n(const float* in, uint8_t* out, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f32_to_e4m3_rne(in[i]);
}

hipError_t launch_bf16_store(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                             uint16_t* Y, uint64_t stride_h, hipStream_t s) {
  uint64_t total = (uint64_t)n * dim;
  uint32_t blocks = (uint32_t)((total + 255) / 256);
  hipLaunchKernelGGL(k_encode_bf16, dim3(blocks), dim3(256), 0, s, X, stride_f, n, dim, Y,
                     stride_h);
  return hipGetLastError();
}

// Per BF16 row x^ (exact in FP32), for every query q.  The search's sums take
// the FP32 query against x^ with FMA, at most 8 NCH + 8 roundings deep:
//  * inner product (metric 1): |<q,x^> - <q,x>| <= |q| |x - x^|, that sum is
//    within gamma |q| |x^| of -<q,x^> and the rerank's float loop within
//    gamma |q| |x| of -<q,x>, so b = |x - x^| + 5.8e-6 g (|x^| + |x|) with
//    |D16 - D32| <= |q| b (gamma <= 5.8e-6 g, g = max(1, dim / 96), as below);
//  * L2 (metric 0): b = |x - x^| + 2^-16 |x^|, as for the FP8 L2 rows below
//    (the difference q_j - x^_j rounds once, the rest is the same sum).
// Computed in double, raised by 1e-6 relative before the float rounding.
__global__ void k_bf16_row_bound(const float* __restrict__ X, uint64_t stride_f, uint32_t n,
                                 uint32_t dim, const uint16_t* __restrict__ Y,
                                 uint64_t stride_h, float* __restrict__ b, int metric) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double rr = 0.0, xn = 0.0, xh = 0.0;
  for (uint32_t j = 0; j < dim; j++) {
    const double x = (double)X[(uint64_t)r * stride_f + j];
    const double h = (double)__uint_as_float((uint32_t)Y[(uint64_t)r * stride_h + j] << 16);
    rr += (x - h) * (x - h);
    xn += x * x;
    xh += h * h;
  }
  const double g = dim > 96u ? (double)dim / 96.0 : 1.0;
  const double v = metric == 0 ? sqrt(rr) + 0x1p-16 * sqrt(xh)
                               : sqrt(rr) + 5.8e-6 * g * (sqrt(xh) + sqrt(xn));
  b[r] = (float)(v * (1.0 + 1e-6));
}

hipError_t launch_bf16_row_bound(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                                 const uint16_t* Y, uint64_t stride_h, float* b, int metric,
                                 hipStream_t s) {
  hipLaunchKernelGGL(k_bf16_row_bound, dim3((n + 255) / 256), dim3(256), 0, s, X, stride_f, n,
                     dim, Y, stride_h, b, metric);
  return hipGetLastError();
}

// Per FP8 row, b >= |D8(q, x) - D32(q, x)| / |q| for every query q, where D8 is the
// search's FP8 inner-product distance and D32 the FP32 rerank's (the reference's
// float loop).  With x^ = s * code (exact reals), r = |x - x^|:
//   |<q,x^> - <q,x>| <= |q| r;
//   D8 rounds q_j s_j to FP32 and then to FP16 after a power-of-two scale 2^e (at
//   most 2^-11 (1 + 2^-12) relative, or 2^-25 2^-e absolute where the FP16 value
//   is subnormal), multiplies exactly and sums in FP32 (at most gamma_96 of the
//   sum of magnitudes, in any order): <= |q| |x^| (4.884e-4 + 5.73e-6) plus
//   96 * 448 * 2^-39 * max|q_j s_j| <= |q| * 7.9e-8 * max_j s_j;
//   D32's float loop of 96 products: <= gamma_97 |q| |x| <= |q| |x| 5.79e-6.
// So b = r + 4.95e-4 |x^| + 5.8e-6 |x| + 8e-8 max_j s_j, computed in double and
// raised by 1e-6 relative before it is rounded to float (dim <= 96 per the
// gamma terms; longer rows scale them by dim / 96).
//
// L2 (metric 0): b = |x - x^| + 2^-16 |x^|, raised the same way.  The search's
// sum of (q_j - s_j fp8_j)^2 (the product and the difference rounded to FP32,
// at most 16 NCH + 8 FMA roundings deep, NCH <= 8) is within 2^-16 (|q - x^| +
// |x^|) of |q - x^| after the square root, and ||q - x^| - |q - x|| <= |x - x^|.
__global__ void k_fp8_row_bound(const float* __restrict__ X, uint64_t stride_f, uint32_t n,
                                uint32_t dim, const float* __restrict__ scales,
                                const uint8_t* __restrict__ Y, uint64_t stride_b,
                                float* __restrict__ b, int metric) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double rr = 0.0, xn = 0.0, xh = 0.0, smax = 0.0;
  for (uint32_t j = 0; j < dim; j++) {
    const double s = (double)scales[j];
    const double x = (double)X[(uint64_t)r * stride_f + j];
    const double h = s * (double)fp8e4m3_to_f32(Y[(uint64_t)r * stride_b + j]);
    rr += (x - h) * (x - h);
    xn += x * x;
    xh += h * h;
    smax = s > smax ? s : smax;
  }
  const double g = dim > 96u ? (double)dim / 96.0 : 1.0;
  const double v = metric == 0 ? sqrt(rr) + 0x1p-16 * sqrt(xh)
                               : sqrt(rr) + (4.884e-4 + 5.73e-6 * g) * sqrt(xh) +
                                     5.8e-6 * g * sqrt(xn) + 8e-8 * g * smax;
  b[r] = (float)(v * (1.0 + 1e-6));
}

// The reference's 8-bit inner-product codes (scale s, codes clipped to
// [-127, 127]) give the search the distance -<cq, cx>, an exact integer.  With
// x^ = cx / s and q^ = cq / s (exact reals):
//   <q,x> - <q^,x^> = <q, x - x^> + <q - q^, x^>,
// so |<cq,cx> - s^2 <q,x>| <= |s q| s |x - x^| + |s q - cq| |cx|.  The rerank's
// float loop of dim products (the reference's) differs from -<q,x> by at most
// gamma_(dim+1) |q| |x| <= 5.8e-6 g |q| |x|, g = max(1, dim / 96).  Per row:
// R = s (|x - x^| + 5.8e-6 g |x|) and N = |cx|, in double, raised by 1e-6
// relative before they are rounded to float.
__global__ void k_q8_row_bound(const float* __restrict__ X, uint64_t stride_f, uint32_t n,
                               uint32_t dim, float scale, const int8_t* __restrict__ Y,
                               uint64_t stride_b, float* __restrict__ b) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double s = (double)scale;
  double rr = 0.0, xn = 0.0, cn = 0.0;
  for (uint32_t j = 0; j < dim; j++) {
    const double x = (double)X[(uint64_t)r * stride_f + j];
    const double c = (double)Y[(uint64_t)r * stride_b + j];
    rr += (x - c / s) * (x - c / s);
    xn += x * x;
    cn += c * c;
  }
  const double g = dim > 96u ? (double)dim / 96.0 : 1.0;
  b[2 * (uint64_t)r] = (float)(s * (sqrt(rr) + 5.8e-6 * g * sqrt(xn)) * (1.0 + 1e-6));
  b[2 * (uint64_t)r + 1] = (float)(sqrt(cn) * (1.0 + 1e-6));
}

hipError_t launch_q8_row_bound(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                               float scale, const int8_t* Y, uint64_t stride_b, float* b,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_q8_row_bound, dim3((n + 255) / 256), dim3(256), 0, s, X, stride_f, n,
                     dim, scale, Y, stride_b, b);
  return hipGetLastError();
}

hipError_t launch_fp8_row_bound(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                                const float* scales, const uint8_t* Y, uint64_t stride_b,
                                float* b, int metric, hipStream_t s) {
  hipLaunchKernelGGL(k_fp8_row_bound, dim3((n + 255) / 256), dim3(256), 0, s, X, stride_f, n,
                     dim, scales, Y, stride_b, b, metric);
  return hipGetLastError();
}

hipError_t launch_fp8_store(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                            uint32_t* absmax_bits, float* scales, uint8_t* Y, uint64_t stride_b,
                            hipStream_t s, const float* cent, uint64_t cstride_f,
                            const uint32_t* assign) {
  hipLaunchKernelGGL(k_zero_u32_quant, dim3((dim + 255) / 256), dim3(256), 0, s, absmax_bits, dim);
  hipLaunchKernelGGL(k_absmax_cols, dim3((n + 255) / 256), dim3(128), 0, s, X, stride_f, n, dim,
                     absmax_bits, cent, cstride_f, assign);
  hipLaunchKernelGGL(k_scales_from_absmax, dim3((dim + 255) / 256), dim3(256), 0, s, absmax_bits,
                     dim, scales);
  uint64_t total = (uint64_t)n * dim;
  uint32_t blocks = (uint32_t)((total + 255) / 256);
  hipLaunchKernelGGL(k_encode_fp8, dim3(blocks), dim3(256), 0, s, X, stride_f, n, dim, scales, Y,
                     stride_b, cent, cstride_f, assign);
  if (cent)
    hipLaunchKernelGGL(k_put_centroid_ids, dim3((n + 255) / 256), dim3(256), 0, s, Y, stride_b, n,
                       assign);
  return hipGetLastError();
}

}  // namespace pann

namespace pann {
hipError_t launch_fp8_codec_probe(float* decode_out, const float* enc_in, uint8_t* enc_out,
                                  uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fp8_decode_all, dim3(1), dim3(256), 0, s, decode_out);
  if (n) hipLaunchKernelGGL(k_fp8_encode_n, dim3((n + 255) / 256), dim3(256), 0, s, enc_in, enc_out, n);
  return hipGetLastError();
}
}  // namespace pann

// ---------------------------------------------------------------------------
// The reference's 8-bit scalar quantization of float points (the quant path of
// python/graph_index.cpp:86-99): uint8 affine codes for L2
// (Euclidian_Point<uint8_t>, euclidian_point.h:98-107, 188-243) and int8
// symmetric codes for MIPS (Quantized_Mips_Point<8, true>, mips_point.h:292-470).
// The parameters are computed on the host from per-row extrema.
// ---------------------------------------------------------------------------
namespace pann {

// per row: min(0, min_j x) and max(0, max_j x); notint |= some x is negative
// or not an integer.  One wave per row, rows grid-strided.
__global__ __launch_bounds__(256) void k_row_minmax(const float* __restrict__ X,
                                                    uint64_t stride_f, uint32_t n, uint32_t dim,
                                                    float* mins, float* maxs, uint32_t* notint) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  bool bad = false;
  for (uint32_t r = wid; r < n; r += nw) {
    float mn = 0.0f, mx = 0.0f;
    const float* x = X + (uint64_t)r * stride_f;
    for (uint32_t j = lane; j < dim; j += 64) {
      const float v = x[j];
      mn = fminf(mn, v);
      mx = fmaxf(mx, v);
      bad = bad || !(v >= 0.0f) || (v - truncf(v)) != 0.0f;
    }
    for (int o = 32; o >= 1; o >>= 1) {
      mn = fminf(mn, __shfl_xor(mn, o));
      mx = fmaxf(mx, __shfl_xor(mx, o));
    }
    if (lane == 0) {
      mins[r] = mn;
      maxs[r] = mx;
    }
  }
  if (bad) atomicOr(notint, 1u);
}

// codes of nrows float rows; metric 0: uint8 affine, 1: int8 symmetric
__global__ void k_encode_q8(const float* __restrict__ X, uint64_t stride_f, uint32_t nrows,
                            uint32_t dim, int metric, float slope, int32_t offset, float max_val,
                            float scale, uint8_t* __restrict__ out, uint64_t out_stride) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)nrows * dim;
  if (i >= total) return;
  const uint64_t r = i / dim, j = i % dim;
  const float x = X[r * stride_f + j];
  uint8_t c;
  if (metric == 0) {
    if (slope == 1.0f && offset == 0) {
      c = (uint8_t)(int)x;  // translate_point's direct copy (euclidian_point.h:193-195)
    } else {
      int64_t v = (int64_t)roundf(x * slope) - offset;
      v = v < 0 ? 0 : (v > 255 ? 255 : v);
      c = (uint8_t)v;
    }
  } else {
    int32_t v = x < -max_val ? -127 : (x > max_val ? 127 : (int32_t)roundf(x * scale));
    c = (uint8_t)(int8_t)v;
  }
  out[r * out_stride + j] = c;
}

hipError_t launch_row_minmax(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                             float* mins, float* maxs, uint32_t* notint, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32_quant, dim3(1), dim3(64), 0, s, notint, 1u);
  uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)n + 3) / 4, 65536);
  hipLaunchKernelGGL(k_row_minmax, dim3(blocks), dim3(256), 0, s, X, stride_f, n, dim, mins,
                     maxs, notint);
  return hipGetLastError();
}

hipError_t launch_encode_q8(const float* X, uint64_t stride_f, uint32_t nrows, uint32_t dim,
                            int metric, float slope, int32_t offset, float max_val, float scale,
                            uint8_t* out, uint64_t out_stride, hipStream_t s) {
  const uint64_t total = (uint64_t)nrows * dim;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_encode_q8, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s, X,
                     stride_f, nrows, dim, metric, slope, offset, max_val, scale, out,
                     out_stride);
  return hipGetLastError();
}

}  // namespace pann
