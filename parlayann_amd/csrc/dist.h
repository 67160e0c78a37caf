// dist.h -- distance policies of the beam-search gather (per-hop GEMV work).
//
// Two evaluation shapes:
//  * group: 8 lanes per row, each lane owns 16-byte chunks at byte offsets
//    sub*16 + c*128 (c < NCH), so every 8 lanes read one full 128-byte line
//    per instruction; partial sums are reduced with DPP inside the 8-lane
//    group.  Integer (uint8/int8) sums are exact in any order, so group mode is
//    bit-exact for them (euclidian_point.h:45-81, mips_point.h:44-58).
//  * in-order: terms summed strictly in index order, the arithmetic of the
//    reference's float loop (euclidian_point.h:83-90, mips_point.h:60-66) as
//    g++ -O3 -march=<FMA target> compiles it: separately rounded products
//    added one at a time up to the last multiple of 4 terms, the last d % 4
//    terms as fused multiply-adds.  So FP32 results are bit-identical to
//    it.  Rows of up to 8 lines keep the group gather and carry the running
//    sum from lane to lane; longer rows go 8 lines at a time with the query
//    in LDS.  One lane per row remains for the rerank of the FP32 store's
//    kernels and of the FP8 L2 kernels.
#pragma once
#include "device_common.h"

namespace pann {

enum StoreKind : int { ST_F32 = 0, ST_U8 = 1, ST_I8 = 2, ST_BF16 = 3, ST_FP8 = 4, ST_FP8R = 5 };
enum MetricKind : int { M_L2 = 0, M_IP = 1 };

__host__ __device__ constexpr int store_elem_bytes(int st) {
  return st == ST_F32 ? 4 : (st == ST_BF16 ? 2 : 1);
}

// ---------------------------------------------------------------------------
// OCP FP8 E4M3 (e4m3fn: bias 7, no infinities, max finite 448) to f32.
// Exact: every e4m3 value is representable in f32.  Normal numbers are
// rebuilt by moving the exponent/mantissa fields into f32 position and
// re-biasing (+120); subnormals (exponent field 0) are m * 2^-9.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fp8e4m3_to_f32(uint32_t x) {
  uint32_t mag = x & 0x7Fu;
  uint32_t sgn = (x & 0x80u) << 24;
  float v;
  if (mag >= 8u) v = __uint_as_float(((mag << 20) + (120u << 23)));
  else v = (float)mag * 0.001953125f;  // m * 2^-9
  return __uint_as_float(__float_as_uint(v) | sgn);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf16lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

// ---------------------------------------------------------------------------
// Group (8 lanes per row) policies.  Q holds this lane's slice of the query.
// ---------------------------------------------------------------------------
template <int ST, int MT, int NCH>
struct GroupDist;

// FP32 rows: 4 floats per 16-byte chunk.
template <int MT, int NCH>
struct GroupDist<ST_F32, MT, NCH> {
  float q[NCH][4];
  template <int LW = 4>  // whole 16-byte slices only
  __device__ __forceinline__ void load_query(const float* qv, const float* /*scales*/, int dim, int sub) {
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        int j = (c * 128 + sub * 16) / 4 + e;
        q[c][e] = j < dim ? qv[j] : 0.0f;
      }
  }
  __device__ __forceinline__ float partial(const uint4* raw) const {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      float p[4] = {__uint_as_float(raw[c].x), __uint_as_float(raw[c].y),
                    __uint_as_float(raw[c].z), __uint_as_float(raw[c].w)};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        if (MT == M_L2) {
          float t = q[c][e] - p[e];
          acc = fmaf(t, t, acc);
        } else {
          acc = fmaf(q[c][e], p[e], acc);
        }
      }
    }
    return acc;
  }
  __device__ __forceinline__ float finish(float s) const { return MT == M_L2 ? s : -s; }
};

// BF16 rows: 8 values per chunk, f32 arithmetic.
template <int MT, int NCH>
struct GroupDist<ST_BF16, MT, NCH> {
  float q[NCH][8];
  template <int LW = 4>  // whole 16-byte slices only
  __device__ __forceinline__ void load_query(const float* qv, const float* /*scales*/, int dim, int sub) {
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        int j = (c * 128 + sub * 16) / 2 + e;
        q[c][e] = j < dim ? qv[j] : 0.0f;
      }
  }
  __device__ __forceinline__ float partial(const uint4* raw) const {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      uint32_t w[4] = {raw[c].x, raw[c].y, raw[c].z, raw[c].w};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float p0 = bf16lo(w[e]), p1 = bf16hi(w[e]);
        if (MT == M_L2) {
          float t0 = q[c][2 * e] - p0, t1 = q[c][2 * e + 1] - p1;
          acc = fmaf(t0, t0, acc);
          acc = fmaf(t1, t1, acc);
        } else {
          acc = fmaf(q[c][2 * e], p0, acc);
          acc = fmaf(q[c][2 * e + 1], p1, acc);
        }
      }
    }
    return acc;
  }
  __device__ __forceinline__ float finish(float s) const { return MT == M_L2 ? s : -s; }
};

// FP8 e4m3 rows with per-dimension scales s_j: x_j ~ s_j * fp8_j.  This form
// serves L2, which keeps the scale beside the query; inner product has the
// FP16 form below, which folds it into the query (q~_j = q_j * s_j).
template <int MT, int NCH>
struct GroupDist<ST_FP8, MT, NCH> {
  float q[NCH][16];
  float s[MT == M_L2 ? NCH : 1][16];
  template <int LW = 4>
  __device__ __forceinline__ void load_query(const float* qv, const float* scales, int dim, int sub) {
    constexpr int lw = LW;
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
      for (int e = 0; e < 16; e++) {
        int j = e < 4 * lw ? c * 128 + sub * 4 * lw + e : dim;
        float qq = j < dim ? qv[j] : 0.0f;
        float sc = j < dim ? scales[j] : 0.0f;
        if (MT == M_L2) {
          q[c][e] = qq;
          s[c][e] = sc;
        } else {
          q[c][e] = qq * sc;
        }
      }
  }
  __device__ __forceinline__ float partial(const uint4* raw) const { return partial_w<4>(raw); }
  template <int LW>  // the lane's first LW words of each chunk
  __device__ __forceinline__ float partial_w(const uint4* raw) const {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      uint32_t w[4] = {raw[c].x, raw[c].y, raw[c].z, raw[c].w};
      float pv[16];
#pragma unroll
      for (int e = 0; e < LW; e++) {  // v_cvt_pk_f32_fp8: two e4m3 bytes per instruction
        f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[e], false);
        f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8((int)w[e], true);
        pv[4 * e + 0] = lo[0];
        pv[4 * e + 1] = lo[1];
        pv[4 * e + 2] = hi[0];
        pv[4 * e + 3] = hi[1];
      }
#pragma unroll
      for (int e = 0; e < 4 * LW; e++) {
        float p = pv[e];
        if (MT == M_L2) {
          float t = q[c][e] - s[c][e] * p;
          acc = fmaf(t, t, acc);
        } else {
          acc = fmaf(q[c][e], p, acc);
        }
      }
    }
    return acc;
  }
  __device__ __forceinline__ float finish(float v) const { return MT == M_L2 ? v : -v; }
};

// FP8 e4m3 rows, inner product.  The scaled query q~_j = q_j * s_j (FP32, as
// above) is multiplied by 2^e, e chosen per query so that max_j |q~_j| * 2^e
// lies in [2^14, 2^15), and rounded to FP16 (RNE): a relative error of at
// most 2^-11 per element, against e4m3's 2^-4 on the row.  Each pair of row
// codes becomes an FP16 pair (v_cvt_scalef32_pk_f16_fp8; exact, every e4m3
// value is an FP16 normal), and v_dot2_f32_f16 adds the two products, exact
// in FP32 (11 x 4 significant bits), to the FP32 sum.  finish() multiplies by
// 2^-e, which is exact.  Half the query registers of the FP32 form above, and
// 16 instructions per 16-byte chunk instead of 24.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float group8_max(float v) {
  v = fmaxf(v, __int_as_float(dpp_mov<0xB1>(__float_as_int(v))));
  v = fmaxf(v, __int_as_float(dpp_mov<0x4E>(__float_as_int(v))));
  v = fmaxf(v, __int_as_float(dpp_mov<0x141>(__float_as_int(v))));
  return v;
}

// The per-query exponent e above for max |q~| = mx (0 for mx = 0), clamped so
// that 2^e and 2^-e stay normal floats.
__device__ __forceinline__ int fp16_query_exp(float mx) {
  if (!(mx > 0.0f)) return 0;
  int e = 15 - __builtin_amdgcn_frexp_expf(mx);  // mx in [2^(E-1), 2^E)
  return e < -120 ? -120 : (e > 120 ? 120 : e);
}

template <int NCH>
struct GroupDist<ST_FP8, M_IP, NCH> {
  f16x2 q[NCH][8];
  float scl, inv;  // 2^e, 2^-e
  template <int LW = 4>
  __device__ __forceinline__ void load_query(const float* qv, const float* scales, int dim, int sub) {
    constexpr int lw = LW;
    float v[NCH][16];
    float mx = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
      for (int e = 0; e < 16; e++) {
        int j = e < 4 * lw ? c * 128 + sub * 4 * lw + e : dim;
        v[c][e] = j < dim ? qv[j] * scales[j] : 0.0f;
        mx = fmaxf(mx, fabsf(v[c][e]));
      }
    // the 8 lanes of a group hold the whole query between them
    const int ex = fp16_query_exp(group8_max(mx));
    scl = __builtin_ldexpf(1.0f, ex);
    inv = __builtin_ldexpf(1.0f, -ex);
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
      for (int e = 0; e < 8; e++) {
        f16x2 h = {(_Float16)(v[c][2 * e] * scl), (_Float16)(v[c][2 * e + 1] * scl)};
        q[c][e] = h;
      }
  }
  __device__ __forceinline__ float partial(const uint4* raw) const { return partial_w<4>(raw); }
  template <int LW>  // the lane's first LW words of each chunk
  __device__ __forceinline__ float partial_w(const uint4* raw) const {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      uint32_t w[4] = {raw[c].x, raw[c].y, raw[c].z, raw[c].w};
#pragma unroll
      for (int e = 0; e < LW; e++) {
        f16x2 lo = __builtin_amdgcn_cvt_scalef32_pk_f16_fp8(w[e], 1.0f, false);
        f16x2 hi = __builtin_amdgcn_cvt_scalef32_pk_f16_fp8(w[e], 1.0f, true);
        acc = __builtin_amdgcn_fdot2(lo, q[c][2 * e], acc, false);
        acc = __builtin_amdgcn_fdot2(hi, q[c][2 * e + 1], acc, false);
      }
    }
    return acc;
  }
  __device__ __forceinline__ float finish(float v) const { return -(v * inv); }
};

// FP8 residual rows (inner product only): x ~ c + s * fp8, c the row's
// centroid, named by the row's last 4 bytes (four 6-bit digits, finite e4m3
// codes that meet a zero query).  -q.x = -(q.c + sum_j q_j s_j fp8_j): the
// lane holding the id word (sub 7 of the last chunk) adds q.c from the
// query's centroid table T (LDS, filled before the search).
__device__ __forceinline__ uint32_t centroid_of(uint32_t w) {
  return (w & 63u) | ((w >> 8) & 63u) << 6 | ((w >> 16) & 63u) << 12 | ((w >> 24) & 63u) << 18;
}

template <int MT, int NCH>
struct GroupDist<ST_FP8R, MT, NCH> : GroupDist<ST_FP8, M_IP, NCH> {
  static_assert(MT == M_IP, "the FP8 residual store is inner product only");
  const float* T;
  bool idl;
  template <int LW = 4>  // whole 16-byte slices only
  __device__ __forceinline__ void load_query(const float* qv, const float* scales, int dim, int sub) {
    GroupDist<ST_FP8, M_IP, NCH>::load_query(qv, scales, dim, sub);
    idl = sub == 7;
  }
  __device__ __forceinline__ float partial(const uint4* raw) const {
    float acc = GroupDist<ST_FP8, M_IP, NCH>::partial(raw);
    // the sum is in units of 2^-e; so is q.c (a power-of-two product, exact)
    if (idl) acc += T[centroid_of(raw[NCH - 1].w)] * this->scl;
    return acc;
  }
};

// uint8 rows: exact int32 via v_dot4_u32_u8.
//   L2: sum (q-p)^2 = sum q^2 + sum p^2 - 2 sum q.p   (mod 2^32, exact result)
template <int MT, int NCH>
struct GroupDist<ST_U8, MT, NCH> {
  uint32_t q[NCH][4];
  uint32_t qq;  // this lane's share of sum q^2
  template <int LW = 4>
  __device__ __forceinline__ void load_query(const uint8_t* qv, const float*, int dim, int sub) {
    constexpr int lw = LW;
    qq = 0u;
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
      for (int w = 0; w < 4; w++) {
        uint32_t v = 0u;
#pragma unroll
        for (int b = 0; b < 4; b++) {
          int j = w < lw ? c * 128 + sub * 4 * lw + 4 * w + b : dim;
          uint32_t x = j < dim ? (uint32_t)qv[j] : 0u;
          v |= x << (8 * b);
        }
        q[c][w] = v;
        qq = __builtin_amdgcn_udot4(v, v, qq, false);
      }
  }
  __device__ __forceinline__ int partial(const uint4* raw) const { return partial_w<4>(raw); }
  template <int LW>  // the lane's first LW words of each chunk
  __device__ __forceinline__ int partial_w(const uint4* raw) const {
    uint32_t pp = MT == M_L2 ? qq : 0u, pq = 0u;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      uint32_t w[4] = {raw[c].x, raw[c].y, raw[c].z, raw[c].w};
#pragma unroll
      for (int e = 0; e < LW; e++) {
        if (MT == M_L2) pp = __builtin_amdgcn_udot4(w[e], w[e], pp, false);
        pq = __builtin_amdgcn_udot4(w[e], q[c][e], pq, false);
      }
    }
    return MT == M_L2 ? (int)(pp - 2u * pq) : (int)pq;
  }
  __device__ __forceinline__ float finish(int s) const { return MT == M_L2 ? (float)s : -((float)s); }
};

// int8 rows: exact int32 via v_dot4_i32_i8.
template <int MT, int NCH>
struct GroupDist<ST_I8, MT, NCH> {
  int32_t q[NCH][4];
  int32_t qq;
  template <int LW = 4>
  __device__ __forceinline__ void load_query(const int8_t* qv, const float*, int dim, int sub) {
    constexpr int lw = LW;
    qq = 0;
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
      for (int w = 0; w < 4; w++) {
        uint32_t v = 0u;
#pragma unroll
        for (int b = 0; b < 4; b++) {
          int j = w < lw ? c * 128 + sub * 4 * lw + 4 * w + b : dim;
          uint32_t x = j < dim ? (uint32_t)(uint8_t)qv[j] : 0u;
          v |= x << (8 * b);
        }
        q[c][w] = (int32_t)v;
        qq = __builtin_amdgcn_sdot4((int)v, (int)v, qq, false);
      }
  }
  __device__ __forceinline__ int partial(const uint4* raw) const { return partial_w<4>(raw); }
  template <int LW>  // the lane's first LW words of each chunk
  __device__ __forceinline__ int partial_w(const uint4* raw) const {
    int pp = MT == M_L2 ? qq : 0, pq = 0;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      int w[4] = {(int)raw[c].x, (int)raw[c].y, (int)raw[c].z, (int)raw[c].w};
#pragma unroll
      for (int e = 0; e < LW; e++) {
        if (MT == M_L2) pp = __builtin_amdgcn_sdot4(w[e], w[e], pp, false);
        pq = __builtin_amdgcn_sdot4(w[e], q[c][e], pq, false);
      }
    }
    return MT == M_L2 ? (int)((uint32_t)pp - 2u * (uint32_t)pq) : pq;
  }
  __device__ __forceinline__ float finish(int s) const { return MT == M_L2 ? (float)s : -((float)s); }
};

// Whether a store's one-chunk rows may run on 12-byte lane slices: 1-byte
// elements with no per-row extra word (the FP8 residual rows end in one).
template <int ST, int NCH>
__host__ __device__ constexpr bool lane_slices_ok() {
  return NCH == 1 && (ST == ST_FP8 || ST == ST_U8 || ST == ST_I8);
}

// Loads the query in the lane slices the gather will read (lw = 3 or 4 words
// per lane and chunk); two straight-line loads, so the slice width stays a
// compile-time constant in each.
template <int ST, int MT, int NCH, typename QT>
__device__ __forceinline__ void load_query_slices(GroupDist<ST, MT, NCH>& D, const QT* qv,
                                                  const float* scales, int dim, int sub, int lw) {
  if constexpr (lane_slices_ok<ST, NCH>()) {
    if (lw == 3) {
      D.template load_query<3>(qv, scales, dim, sub);
      return;
    }
  }
  D.template load_query<4>(qv, scales, dim, sub);
}

template <int ST>
struct QueryElem {
  using T = float;
};
template <>
struct QueryElem<ST_U8> {
  using T = uint8_t;
};
template <>
struct QueryElem<ST_I8> {
  using T = int8_t;
};

// DPP moves inside each group of 8 lanes: the mirror (lane i takes lane
// 7 - i), xor 2 and xor 1 -- each pairs a lane with one whose bit 2, 1 or 0
// differs.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_mov<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ int dpp_f(int v) {
  return dpp_mov<CTRL>(v);
}

// One step of the transpose reduction: each lane holds N partial sums (rows
// 0..N-1 of its current set) and a partner across bit `b`; it keeps the
// half of the rows whose index bit equals its own bit, adding the partner's
// partials of those rows (the partner sends the half it does not keep).
// With N = 1 both lanes keep the one row.
template <int CTRL, int N, typename T>
__device__ __forceinline__ void transpose_step(T (&p)[N], T (&q)[N > 1 ? N / 2 : 1], bool b) {
  if constexpr (N == 1) {
    q[0] = p[0] + dpp_f<CTRL>(p[0]);
  } else {
#pragma unroll
    for (int j = 0; j < N / 2; j++) {
      const T send = b ? p[j] : p[j + N / 2];
      const T keep = b ? p[j + N / 2] : p[j];
      q[j] = keep + dpp_f<CTRL>(send);
    }
  }
}

// The row sums of a group's UP rows from the 8 lanes' partials p[0..UP):
// three exchange-and-add steps across lane bits 2, 1 and 0 (7 DPP adds for
// UP = 8 instead of 3 per row).  Lane `sub` ends with the sum of row
// sub >> (3 - log2 UP); for UP < 8 two or four lanes hold each row.
template <int UP, typename T>
__device__ __forceinline__ T group8_transpose_sum(T (&p)[UP], int sub) {
  constexpr int U1 = UP > 1 ? UP / 2 : 1;
  constexpr int U2 = U1 > 1 ? U1 / 2 : 1;
  T q1[U1], q2[U2], q3[1];
  transpose_step<0x141, UP>(p, q1, (sub & 4) != 0);
  transpose_step<0x4E, U1>(q1, q2, (sub & 2) != 0);
  transpose_step<0xB1, U2>(q2, q3, (sub & 1) != 0);
  return q3[0];
}

// Evaluates the distances of rows NL[0..ncmp) into DL[0..ncmp) (both LDS).
// Rows are gathered 8 at a time (one per 8-lane group), UP rounds of loads in
// flight per lane.  A row's address is the lane's base plus (id * lines) << 7,
// lines = the row's 128-byte lines: the product fits 32 bits for any index
// under 512 GB, and one 64-bit shift-and-add (v_lshl_add_u64) places it.
// PAD: NL is this wave's own and holds at least round_up(ncmp, 8 * UP)
// entries; those past ncmp are set to NL[0] first, so each round's slots
// read ids without a bounds select (their distances are discarded).
// LW: 4-byte words per lane and chunk; 3 for one-chunk rows of at most 96
// bytes, whose 8 lanes then read 12 bytes each and leave the line's padding
// unread (the query was loaded with the same slices, load_query's lw).
template <int ST, int MT, int NCH, bool PAD, int LW>
__device__ __forceinline__ void eval_group_t(const GroupDist<ST, MT, NCH>& D, const char* base,
                                             uint64_t stride, int row_bytes, uint32_t* NL,
                                             int ncmp, float* DL, int lane) {
  // rows in flight per lane: every 8-lane group gathers UP rows per round,
  // sized so the raw chunks (UP * NCH * 4 VGPRs) keep the kernel at or
  // under ~128 VGPRs, i.e. no fewer waves per CU than LDS already allows
  // (FP8 and BF16 rows keep UP * NCH <= 8: their query slices take 8 or 16
  // registers per chunk, which leaves room for only that many raw chunks)
  constexpr bool wideq = ST == ST_FP8 || ST == ST_FP8R || ST == ST_BF16;
  constexpr int UP = wideq && NCH >= 2 && NCH <= 4 ? (NCH == 2 ? 4 : 2)
                     : NCH <= 2 ? 8 : (NCH <= 4 ? 4 : 2);
  constexpr int RSH = UP == 8 ? 0 : (UP == 4 ? 1 : 2);  // lane sub holds row sub >> RSH
  const int sub = lane & 7, grp = lane >> 3;
  // 128-byte lines per row: the chunk count itself up to 4 (the stride is the
  // row rounded up to 128 bytes), so the multiply folds into shifts and adds
  const uint32_t lines = NCH <= 4 ? (uint32_t)NCH : (uint32_t)(stride >> 7);
  if (ncmp <= 0) return;
  const char* lbase = base + sub * 4 * LW;  // this lane's slice of every row
  if constexpr (PAD) {
    const int np = (ncmp + 8 * UP - 1) / (8 * UP) * (8 * UP);
    const uint32_t id0 = NL[0];
    for (int r = ncmp + lane; r < np; r += PANN_WAVE) NL[r] = id0;
    wave_lds_sync();
  }
  // 128-byte lines a row occupies: NCH rounds 5..7 up to 8, and the chunks
  // past the row's last line are not loaded (they would be the next row's
  // line, or past the last row's end)
  const int nl = (row_bytes + 127) >> 7;
  // Branch-free gathers: slots past ncmp re-read row NL[0] (an L1 hit) and are
  // discarded; bytes past row_bytes within its last line read the row's zero
  // padding (rows are padded to a 128-byte multiple, the query is zero
  // there), which adds nothing to either metric and no cache lines.
  for (int r0 = 0; r0 < ncmp; r0 += 8 * UP) {
    uint4 raw[UP][NCH];
#pragma unroll
    for (int u = 0; u < UP; u++) {
      const int r = r0 + u * 8 + grp;
      const uint32_t id = PAD ? NL[r] : NL[r < ncmp ? r : 0];
      const char* row = lbase + ((uint64_t)(id * lines) << 7);
      if constexpr (LW == 3) {
        const uint3 v = *(const uint3*)row;
        raw[u][0] = make_uint4(v.x, v.y, v.z, 0u);
      } else {
#pragma unroll
        for (int c = 0; c < NCH; c++) {
          if (NCH <= 4 || c < nl)  // uniform: the same for every row of the launch
            raw[u][c] = ld16(row + c * 128);
          else
            raw[u][c] = make_uint4(0u, 0u, 0u, 0u);
        }
      }
    }
    using PT = decltype(D.partial(raw[0]));
    PT part[UP];
#pragma unroll
    for (int u = 0; u < UP; u++) {
      if constexpr (LW == 4) part[u] = D.partial(raw[u]);
      else part[u] = D.template partial_w<LW>(raw[u]);
    }
    const PT tot = group8_transpose_sum<UP>(part, sub);
    const int r = r0 + (sub >> RSH) * 8 + grp;
    if ((sub & ((1 << RSH) - 1)) == 0 && r < ncmp) DL[r] = D.finish(tot);
  }
}


template <int ST, int MT, int NCH, bool PAD>
__device__ __forceinline__ void eval_group(const GroupDist<ST, MT, NCH>& D, const char* base,
                                           uint64_t stride, int row_bytes, int lw, uint32_t* NL,
                                           int ncmp, float* DL, int lane) {
  if constexpr (lane_slices_ok<ST, NCH>()) {
    if (lw == 3) {
      eval_group_t<ST, MT, NCH, PAD, 3>(D, base, stride, row_bytes, NL, ncmp, DL, lane);
      return;
    }
  }
  eval_group_t<ST, MT, NCH, PAD, 4>(D, base, stride, row_bytes, NL, ncmp, DL, lane);
}

// ---------------------------------------------------------------------------
// In-order FP32 over the group layout.  Lane `sub` of a row's 8-lane group
// holds elements c*32 + sub*4 + e (e < 4) of chunk c.  The running sum walks
// the lanes in element order: at step (c, s) lane s takes the sum from lane
// s - 1 (DPP row_shr:1) or, at s = 0, from lane 7 of its own group, which
// ended chunk c - 1 (row_ror:9 maps lane 7 to lane 0 and lane 15 to lane 8 of
// each 16-lane row), and adds its four terms in order.  Lane 7 ends with the
// row's sum.  Every other lane computes a discarded value.  Each term is a
// separately rounded product added with one rounding, except in the group of
// four that holds the last dim % 4 elements (group dim / 4, when dim % 4 != 0),
// whose terms are fused multiply-adds -- the reference loop's compiled
// arithmetic.  The zero padding past dim adds exact zeros to a sum that is
// never -0, so the result is bit-identical to seq_f32_row below and to the
// reference's float loop.
// ---------------------------------------------------------------------------
// The operands (a, b) of the four terms a * b of one 16-byte chunk.
template <int MT>
__device__ __forceinline__ void inorder_ops(const float* q, const uint4& raw, float* a, float* b) {
#pragma clang fp contract(off)
  const float p[4] = {__uint_as_float(raw.x), __uint_as_float(raw.y), __uint_as_float(raw.z),
                      __uint_as_float(raw.w)};
#pragma unroll
  for (int e = 0; e < 4; e++) {
    if (MT == M_L2) {
      a[e] = q[e] - p[e];
      b[e] = a[e];
    } else {
      a[e] = q[e];
      b[e] = p[e];
    }
  }
}

// The fused group of a row of `dim` floats, as (chunk, step): dim / 4 = c * 8
// + s; (-1, -1) when dim % 4 == 0.
struct FusedGroup {
  int c, s;
  __host__ __device__ explicit FusedGroup(int dim)
      : c(dim % 4 ? (dim / 4) / 8 : -1), s(dim % 4 ? (dim / 4) % 8 : -1) {}
};

// The eight steps of one 128-byte chunk; `first` starts the sum at zero.  With
// FUSE, step `fs` (uniform, -1 for none) adds its terms fused: both sums are
// formed and a select keeps one, so the chains of the rows a lane interleaves
// stay in one basic block.  Rows whose dimension is a multiple of 4 run the
// FUSE = false body, which has no fused step at all.
template <bool FUSE>
__device__ __forceinline__ float inorder_chunk(float acc, const float* a, const float* b,
                                               bool first, int fs) {
#pragma clang fp contract(off)
  float t[4];
#pragma unroll
  for (int e = 0; e < 4; e++) t[e] = a[e] * b[e];
#pragma unroll
  for (int s = 0; s < 8; s++) {
    float in;
    if (s == 0) in = first ? 0.0f : __int_as_float(dpp_mov<0x129>(__float_as_int(acc)));
    else in = __int_as_float(dpp_mov<0x111>(__float_as_int(acc)));
    float x = in + t[0];
    x = x + t[1];
    x = x + t[2];
    x = x + t[3];
    if constexpr (FUSE) {
      float y = fmaf(a[0], b[0], in);
      y = fmaf(a[1], b[1], y);
      y = fmaf(a[2], b[2], y);
      y = fmaf(a[3], b[3], y);
      x = s == fs ? y : x;
    }
    acc = x;
  }
  return acc;
}

template <int MT, int NCH, bool FUSE>
__device__ __forceinline__ float group_inorder_sum(const GroupDist<ST_F32, MT, NCH>& D,
                                                   const uint4* raw, FusedGroup fg) {
  float a[NCH][4], b[NCH][4];
#pragma unroll
  for (int c = 0; c < NCH; c++) inorder_ops<MT>(D.q[c], raw[c], a[c], b[c]);
  float acc = 0.0f;
#pragma unroll
  for (int c = 0; c < NCH; c++)
    acc = inorder_chunk<FUSE>(acc, a[c], b[c], c == 0, c == fg.c ? fg.s : -1);
  return MT == M_L2 ? acc : -acc;
}

// The same sum over FP32 rows of any length, the line count known only at run
// time and the query read from LDS (QL, dim floats): the fused rerank, whose
// rows are not the search store's, and rows past 8 lines.  Rows are gathered
// MAXC lines at a time, UP rows per group in flight; the sum carries over from
// one block of lines to the next.
template <int MT, int MAXC, int UP, bool FUSE>
__device__ __forceinline__ void eval_group_inorder_lds_t(const float* QL, int dim,
                                                         const char* base, uint64_t stride,
                                                         const uint32_t* NL, int ncmp, float* DL,
                                                         int lane) {
  const int sub = lane & 7, grp = lane >> 3;
  const int nl = (dim * 4 + 127) >> 7;  // lines per row, taken MAXC at a time
  const FusedGroup fg(dim);
  for (int r0 = 0; r0 < ncmp; r0 += 8 * UP) {
    const char* row[UP];
#pragma unroll
    for (int u = 0; u < UP; u++) {
      const int r = r0 + u * 8 + grp;
      row[u] = base + (uint64_t)NL[r < ncmp ? r : 0] * stride;
    }
    float acc[UP] = {};
    for (int c0 = 0; c0 < nl; c0 += MAXC) {  // uniform
      uint4 raw[UP][MAXC];
#pragma unroll
      for (int u = 0; u < UP; u++)
#pragma unroll
        for (int c = 0; c < MAXC; c++)
          raw[u][c] = c0 + c < nl ? ld16(row[u] + (c0 + c) * 128 + sub * 16)
                                  : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
      for (int c = 0; c < MAXC; c++) {
        if (c0 + c < nl) {  // uniform
          float q[4];
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const int j = (c0 + c) * 32 + sub * 4 + e;
            q[e] = j < dim ? QL[j] : 0.0f;
          }
          const bool first = c0 + c == 0;
          const int fs = c0 + c == fg.c ? fg.s : -1;
#pragma unroll
          for (int u = 0; u < UP; u++) {
            float a[4], b[4];
            inorder_ops<MT>(q, raw[u][c], a, b);
            acc[u] = inorder_chunk<FUSE>(first ? 0.0f : acc[u], a, b, first, fs);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < UP; u++) {
      const int r = r0 + u * 8 + grp;
      if (sub == 7 && r < ncmp) DL[r] = MT == M_L2 ? acc[u] : -acc[u];
    }
  }
}

template <int MT, int MAXC, int UP>
__device__ __forceinline__ void eval_group_inorder_lds(const float* QL, int dim, const char* base,
                                                       uint64_t stride, const uint32_t* NL,
                                                       int ncmp, float* DL, int lane) {
  if (dim % 4)  // uniform
    eval_group_inorder_lds_t<MT, MAXC, UP, true>(QL, dim, base, stride, NL, ncmp, DL, lane);
  else
    eval_group_inorder_lds_t<MT, MAXC, UP, false>(QL, dim, base, stride, NL, ncmp, DL, lane);
}

// eval_group's gather with the in-order sum: DL[r] comes from lane 7 of the
// row's group.
template <int MT, int NCH, bool FUSE>
__device__ __forceinline__ void eval_group_inorder_t(const GroupDist<ST_F32, MT, NCH>& D,
                                                     const char* base, uint64_t stride,
                                                     int row_bytes, const uint32_t* NL,
                                                     int ncmp, float* DL, int lane) {
  constexpr int UP = NCH <= 2 ? 8 : (NCH <= 4 ? 4 : 2);
  const int sub = lane & 7, grp = lane >> 3;
  const int nl = (row_bytes + 127) >> 7;
  const FusedGroup fg(row_bytes >> 2);
  for (int r0 = 0; r0 < ncmp; r0 += 8 * UP) {
    uint4 raw[UP][NCH];
#pragma unroll
    for (int u = 0; u < UP; u++) {
      const int r = r0 + u * 8 + grp;
      const char* row = base + (uint64_t)NL[r < ncmp ? r : 0] * stride;
#pragma unroll
      for (int c = 0; c < NCH; c++) {
        if (NCH <= 4 || c < nl)
          raw[u][c] = ld16(row + c * 128 + sub * 16);
        else
          raw[u][c] = make_uint4(0u, 0u, 0u, 0u);
      }
    }
#pragma unroll
    for (int u = 0; u < UP; u++) {
      const int r = r0 + u * 8 + grp;
      const float s = group_inorder_sum<MT, NCH, FUSE>(D, raw[u], fg);
      if (sub == 7 && r < ncmp) DL[r] = s;
    }
  }
}

template <int MT, int NCH>
__device__ __forceinline__ void eval_group_inorder(const GroupDist<ST_F32, MT, NCH>& D,
                                                   const char* base, uint64_t stride,
                                                   int row_bytes, const uint32_t* NL, int ncmp,
                                                   float* DL, int lane) {
  if (row_bytes % 16)  // uniform: a dimension that is not a multiple of 4
    eval_group_inorder_t<MT, NCH, true>(D, base, stride, row_bytes, NL, ncmp, DL, lane);
  else
    eval_group_inorder_t<MT, NCH, false>(D, base, stride, row_bytes, NL, ncmp, DL, lane);
}

// ---------------------------------------------------------------------------
// Sequential FP32 (one lane per row, in-order sum; the reference's float loop
// as compiled: separately rounded terms up to the last multiple of 4, the
// last dim % 4 terms fused).  The query sits in LDS (QL, dim floats).
// ---------------------------------------------------------------------------
template <int MT>
__device__ __forceinline__ float seq_f32_row(const float* QL, const float* row, int dim) {
  // HIP compiles with fp-contract=fast: without this pragma the multiply and
  // the add below would be fused into one FMA (a single rounding).  Plain
  // operators are used because the pragma does not reach into helper bodies.
#pragma clang fp contract(off)
  float acc = 0.0f;
  int j = 0;
  for (; j + 4 <= dim; j += 4) {
    float4 p = *(const float4*)(row + j);
    float4 q = *(const float4*)(QL + j);
    float t0, t1, t2, t3;
    if (MT == M_L2) {
      t0 = q.x - p.x; t1 = q.y - p.y; t2 = q.z - p.z; t3 = q.w - p.w;
      t0 = t0 * t0; t1 = t1 * t1; t2 = t2 * t2; t3 = t3 * t3;
    } else {
      t0 = q.x * p.x; t1 = q.y * p.y; t2 = q.z * p.z; t3 = q.w * p.w;
    }
    acc = acc + t0;
    acc = acc + t1;
    acc = acc + t2;
    acc = acc + t3;
  }
  for (; j < dim; j++) {  // the last dim % 4 terms, fused
    if (MT == M_L2) {
      const float t = QL[j] - row[j];
      acc = fmaf(t, t, acc);
    } else {
      acc = fmaf(QL[j], row[j], acc);
    }
  }
  return MT == M_L2 ? acc : -acc;
}

template <int MT>
__device__ __forceinline__ void eval_seq_f32(const float* QL, const char* base, uint64_t stride,
                                             int dim, const uint32_t* NL, int ncmp, float* DL,
                                             int lane) {
  for (int r = lane; r < ncmp; r += PANN_WAVE)
    DL[r] = seq_f32_row<MT>(QL, (const float*)(base + (uint64_t)NL[r] * stride), dim);
}

}  // namespace pann
