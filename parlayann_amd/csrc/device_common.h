// device_common.h -- wave64 helpers shared by the CDNA4 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PANN_WAVE 64
#define PANN_SENTINEL 0xFFFFFFFFu

namespace pann {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of `m` strictly below this lane.
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Orders this wave's LDS accesses: every lane's earlier LDS writes are seen by
// every lane's later LDS reads.  A wave's DS instructions execute in issue
// order, so only the compiler has to be kept from moving accesses across.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T readfirstlane(T v) {
  static_assert(sizeof(T) == 4, "32-bit only");
  return __builtin_bit_cast(T, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// ---------------------------------------------------------------------------
// (distance, id) keys: the reference orders pairs by `less` (beamSearch.h:62-64):
// a.dist < b.dist || (a.dist == b.dist && a.id < b.id).  For finite floats the
// map below is monotone, and -0.0 is folded onto +0.0 first (they compare
// equal as floats), so u64 key order == `less` order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f2ord(float d) {
  uint32_t b = __float_as_uint(d);
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
  uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
  return __uint_as_float(b);
}
__device__ __forceinline__ uint64_t make_key(float d, uint32_t id) {
  return ((uint64_t)f2ord(d) << 32) | id;
}
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return (uint32_t)k; }
__device__ __forceinline__ float key_dist(uint64_t k) { return ord2f((uint32_t)(k >> 32)); }

// ---------------------------------------------------------------------------
// DPP reductions inside groups of 8 consecutive lanes (quad_perm xor 1,
// quad_perm xor 2, row_half_mirror).  Every lane of the group ends with the
// same, order-fixed sum.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

__device__ __forceinline__ float group8_sum(float v) {
  v = v + __int_as_float(dpp_mov<0xB1>(__float_as_int(v)));
  v = v + __int_as_float(dpp_mov<0x4E>(__float_as_int(v)));
  v = v + __int_as_float(dpp_mov<0x141>(__float_as_int(v)));
  return v;
}
__device__ __forceinline__ int group8_sum(int v) {
  v = v + dpp_mov<0xB1>(v);
  v = v + dpp_mov<0x4E>(v);
  v = v + dpp_mov<0x141>(v);
  return v;
}

// 16-byte global load through the vector memory path.
__device__ __forceinline__ uint4 ld16(const void* p) { return *(const uint4*)p; }

}  // namespace pann
