// kernels.h -- argument blocks and launchers of the CDNA4 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pann {

// per-launch status bits (OR-ed by the kernels)
enum : uint32_t {
  PANN_ST_SHORT_BEAM = 1u,        // a frontier ended with fewer than k entries
  PANN_ST_HASH_CAPACITY = 2u,     // a query outgrew its visited-set capacity
  PANN_ST_VISITED_OVERFLOW = 4u,  // build mode: visited list longer than vis_cap
  PANN_ST_BAD_PARAMS = 8u,        // a per-query k / beam outside the launch's sizing
};

// An HNSW level searched at beam 1 before the main search (HNSW.hpp:1545-1566).
constexpr uint32_t PANN_MAX_UP = 16;
struct UpperLevel {
  const uint32_t* nbrs;  // n rows of nbr_stride uint32, sentinel past the degree
  uint32_t nbr_stride;
  uint32_t max_degree;
};

struct SearchArgs {
  // search store: n rows of row_bytes valid bytes, pts_stride apart (multiple of 128)
  const char* pts;
  uint64_t pts_stride;
  uint32_t row_bytes;
  // 4-byte words each of a row's 8 lanes reads per 128-byte chunk: 4, or 3
  // for one-chunk rows of 1-byte elements of at most 96 bytes
  uint32_t lane_words;
  const float* scales;  // FP8 per-dimension scales (nullptr otherwise)
  // Rows with the FP32 rerank, per-row error bounds with which the rerank skips
  // rows that cannot reach the first k (nullptr: every row is re-scored):
  //  * FP8, inner product: one float b per row, |FP8 distance - FP32 distance|
  //    <= |q| b for every query q (launch_fp8_row_bound);
  //  * FP8, L2: one float b per row, |sqrt(FP8 distance) - |q - x|| <= b +
  //    2^-16 sqrt(FP8 distance) for every query q (launch_fp8_row_bound);
  //  * BF16: the same two forms (launch_bf16_row_bound);
  //  * the reference's 8-bit codes: two floats (R, N) per row,
  //    |code distance - s^2 FP32 distance| <= |s q| R + |s q - code(q)| N for
  //    every query q, s = q8_scale (launch_q8_row_bound)
  const float* row_bound;
  float q8_scale;
  // FP8 residual store: q . c for every query and centroid (nq x ncent,
  // computed before the launch); each row's last 4 bytes name its centroid
  const float* qtable;
  uint32_t ncent;
  // FP32 store for the rerank branch (nullptr: plain branch)
  const char* full;
  uint64_t full_stride;
  // graph: n rows of nbr_stride uint32, entries past the degree = sentinel
  const uint32_t* nbrs;
  uint32_t nbr_stride;
  uint32_t max_degree;
  uint32_t n, dim;
  // queries (search format) and FP32 queries for the rerank
  const char* queries;
  uint64_t q_stride;
  const char* fqueries;
  uint64_t fq_stride;
  uint32_t nq;
  // start points (device) shared by all queries, 1 <= nstarts <= beam ...
  const uint32_t* starts;
  uint32_t nstarts;
  // ... or per query: starts_pq[q * starts_stride .. + nstarts_pq[q]) (nstarts_pq
  // NULL: one start each), 1 <= count <= the query's beam
  const uint32_t* starts_pq;
  const uint32_t* nstarts_pq;
  uint32_t starts_stride;
  // 1: add this launch's visits and evaluations to out_visited / out_cmps
  uint32_t accumulate;
  // 1: a frontier shorter than k is a result, not PANN_ST_SHORT_BEAM
  uint32_t short_ok;
  // 1: while a merge is deferred, gather the next visit's rows in the same
  // round (two visits per round); 2: the same, but only in rounds where the
  // merge rule applied to a prediction of the pending candidates defers;
  // 0: one visit per round
  uint32_t lookahead;
  // QueryParams (types.h:233-268); per-query k / beam override the uniform ones
  int32_t k, beam;
  const int32_t* k_pq;
  const int32_t* beam_pq;
  // nullptr, or the order in which the persistent grid takes the queries (the
  // longest beams first, launch_order_by_beam); results do not depend on it
  const uint32_t* qorder;
  int64_t limit;
  int32_t degree_limit;
  double batch_factor;
  double rerank_factor;
  // build mode: the query's own id (same_as) and the visited list (u64 keys)
  const uint32_t* self_ids;
  uint64_t* vis_out;
  uint32_t vis_cap;
  // outputs
  uint32_t* out_ids;
  float* out_dists;
  uint32_t out_stride;
  uint32_t* out_visited;
  uint32_t* out_cmps;
  uint32_t* out_nfront;
  uint32_t* status;
  // scratch
  uint32_t* work_counter;
  unsigned long long* ghash;  // grid x 2^ghash_log2 slots of (tag << 32 | id)
  uint32_t ghash_log2;
  uint32_t ghash_tag_base;    // pass p of query q uses tag base + q * (nup + 1) + p + 1
  // 1: the seen set holds every id at first encounter (dist_cmps equal to the
  // reference's); 0: it holds accepted ids only (same results, fewer inserts)
  uint32_t exact_seen;
  // LDS layout
  uint32_t fcap, ccap, hslots_log2;
  // log2 of the slots of the LDS cache of evaluated ids (0 = none): a
  // direct-mapped table, a slot holding the last committed id that hashed there
  uint32_t rc_log2;
  uint32_t fb_log2;         // bits of the frontier-id Bloom filter (log2), 0 = none
  // 1: the FP32 query row has an LDS region of its own for the whole search (the
  // in-order sum of rows past 8 lines, or a row too long for
  // the candidate buffers); 0: the rerank stages it in the candidate buffers,
  // which are dead by then, and no other step reads it from LDS
  uint32_t q_own;
  // HNSW descent in the same launch: before the search above, each query is
  // searched at beam 1 on up[0], up[1], .. up[nup - 1] (top level first), the
  // first pass from the start points, each later pass and the main search
  // from the previous pass's nearest id; visits and evaluations add up
  uint32_t nup;
  UpperLevel up[PANN_MAX_UP];
  // waves per query (1, 2 or 4): a workgroup of wpq waves takes each query,
  // wave 0 running the search and every wave sharing its gathers, merges and
  // rerank (the launch's block size is 64 * wpq)
  uint32_t wpq;
};

uint32_t search_lds_bytes(const SearchArgs& a);
// per-store launchers (bs_*.hip), called by launch_beam_search
hipError_t launch_bs_f32(const SearchArgs& a, int metric, int nch, uint32_t grid, hipStream_t s);
hipError_t launch_bs_f32_inorder(const SearchArgs& a, int metric, int nch, uint32_t grid,
                                 hipStream_t s);
hipError_t launch_bs_u8(const SearchArgs& a, int metric, int nch, uint32_t grid, hipStream_t s);
hipError_t launch_bs_i8(const SearchArgs& a, int metric, int nch, uint32_t grid, hipStream_t s);
hipError_t launch_bs_bf16(const SearchArgs& a, int metric, int nch, uint32_t grid, hipStream_t s);
hipError_t launch_bs_fp8(const SearchArgs& a, int metric, int nch, uint32_t grid, hipStream_t s);
hipError_t launch_bs_fp8r(const SearchArgs& a, int metric, int nch, uint32_t grid, hipStream_t s);
// the search kernel; a.work_counter and a.status must be zero at its start (the
// library's launch set-up clears both in one kernel before each search)
hipError_t launch_beam_search(const SearchArgs& a, int store, int metric, bool sequential,
                              uint32_t grid, hipStream_t s);

}  // namespace pann

namespace pann {
hipError_t launch_bf16_store(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                             uint16_t* Y, uint64_t stride_h, hipStream_t s);
// cent != nullptr: the residual store, x - cent[assign[i]] is encoded and the
// row's last 4 bytes (stride_b - 4) hold assign[i] (four 6-bit digits)
hipError_t launch_fp8_store(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                            uint32_t* absmax_bits, float* scales, uint8_t* Y, uint64_t stride_b,
                            hipStream_t s, const float* cent = nullptr, uint64_t cstride_f = 0,
                            const uint32_t* assign = nullptr);
// b[0 .. n): per FP8 row (the store of launch_fp8_store, uncentred), over every
// query q: inner product (metric 1), an upper bound on |FP8 distance - FP32
// distance| / |q|; L2 (metric 0), one on |sqrt(FP8 distance) - |q - x|| less
// 2^-16 sqrt(FP8 distance)
hipError_t launch_fp8_row_bound(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                                const float* scales, const uint8_t* Y, uint64_t stride_b,
                                float* b, int metric, hipStream_t s);
// b[0 .. n): per BF16 row Y, the same bounds as launch_fp8_row_bound's for the
// BF16 distance (inner product: |BF16 distance - FP32 distance| <= |q| b)
hipError_t launch_bf16_row_bound(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                                 const uint16_t* Y, uint64_t stride_h, float* b, int metric,
                                 hipStream_t s);
// b[0 .. 2n): per row of the reference's 8-bit inner-product codes Y (scale s),
// R = s (|x - Y/s| + gamma |x|) and N = |Y|: |-<code(q), Y> - s^2 D32(q, x)| <=
// |s q| R + |s q - code(q)| N for every query q
hipError_t launch_q8_row_bound(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                               float scale, const int8_t* Y, uint64_t stride_b, float* b,
                               hipStream_t s);
// brute-force candidates on the matrix cores (bruteforce.hip)
struct BFArgs {
  const void* X;          // points (f32 / u8 / i8), row stride xs elements
  uint64_t xs;
  const float* xnorm;     // |x|^2 per point (L2), NULL for MIPS
  uint32_t n, dim, dim_pad;  // dim_pad: bf_dim_pad(dim, dtype)
  const float* Q;         // queries, row stride qs floats
  uint64_t qs;
  uint32_t nq;
  int metric;             // 0 L2, 1 MIPS
  uint32_t nranges;       // point ranges per query tile
  uint64_t range_len;
  uint32_t kk, cap;       // kept per query; per-(query, range) buffer, a power of two >= kk + 32
  uint64_t* buf;          // nq x nranges x cap keys
  uint32_t* counts;       // nq x nranges
};
uint32_t bf_lds_bytes(uint32_t cap);
uint32_t bf_dim_pad(uint32_t dim, int point_dtype);  // dim rounded up to 32 bytes
hipError_t launch_bf(const BFArgs& a, int point_dtype, uint32_t* out_ids, float* out_dists,
                     hipStream_t s);
// T (nq x ncent, row-major) = Q (nq rows of qs floats, dim used) . C^T (C: ncent
// rows of cs floats, zero from dim to the next multiple of 8; qs, cs multiples of 4)
hipError_t launch_qc_table(const float* Q, uint64_t qs, uint32_t nq, uint32_t dim,
                           const float* C, uint64_t cs, uint32_t ncent, float* T,
                           hipStream_t s);
hipError_t launch_row_sqnorm(const void* X, int point_dtype, uint64_t xs, uint32_t n,
                             uint32_t dim, float* out, hipStream_t s);

// the reference's 8-bit scalar quantization (per-row extrema; codes of rows)
hipError_t launch_row_minmax(const float* X, uint64_t stride_f, uint32_t n, uint32_t dim,
                             float* mins, float* maxs, uint32_t* notint, hipStream_t s);
hipError_t launch_encode_q8(const float* X, uint64_t stride_f, uint32_t nrows, uint32_t dim,
                            int metric, float slope, int32_t offset, float max_val, float scale,
                            uint8_t* out, uint64_t out_stride, hipStream_t s);
hipError_t launch_fp8_codec_probe(float* decode_out, const float* enc_in, uint8_t* enc_out,
                                  uint32_t n, hipStream_t s);
}  // namespace pann

namespace pann {
struct PruneArgs {
  const char* pts;
  uint64_t stride;
  uint32_t dim;
  const uint32_t* nbrs;
  uint32_t nbr_stride, max_degree;
  const uint32_t* points;  // np point ids
  uint32_t np;
  const uint32_t* cand_ids;
  const float* cand_dists;  // nullptr: compute d(c, p)
  const uint32_t* cand_count;
  uint32_t cand_stride;
  const uint32_t* cand_offsets;  // CSR groups (np + 1 offsets) instead of count/stride
  uint64_t* gkeys;               // per-block global sort scratch (grid x sort_cap)
  uint8_t* gflags;
  double alpha;
  uint32_t R;
  int32_t add;
  uint32_t* out_ids;
  uint32_t* out_count;
  uint32_t sort_cap;  // power of two >= cand_stride + max_degree
  // nullptr, or a counter to which each workgroup adds the number of row
  // distances it evaluated
  unsigned long long* evals;
};

uint32_t prune_lds_bytes(uint32_t sort_cap, uint32_t dim, int dtype);
hipError_t launch_robust_prune(const PruneArgs& a, int dtype, int metric, bool keys_global,
                               uint32_t grid, hipStream_t s);
hipError_t launch_sort_visited(const uint64_t* vis, uint32_t nq, uint32_t cap,
                               const uint32_t* counts, uint32_t* ids, float* dists,
                               hipStream_t s);
// HNSW::search's final sort-and-cut with ties (HNSW.hpp:1590-1596) and the
// node -> point id map
hipError_t launch_hnsw_cut(uint32_t* ids, float* dists, uint32_t stride, uint32_t* count,
                           uint32_t nq, uint32_t k, const int32_t* k_pq,
                           const uint32_t* node_ids, hipStream_t s);
// order[0 .. nq): the queries by descending beam_pq (beams clamped to
// [0, max_beam], max_beam < 8192), a counting sort in one workgroup; the
// order inside one beam value is unspecified.  The same launch clears the
// search launch's status word and work counter (both non-null).
hipError_t launch_order_by_beam(const int32_t* beam_pq, uint32_t nq, uint32_t max_beam,
                                uint32_t* order, uint32_t* status, uint32_t* counter,
                                hipStream_t s);
// acc[0] += v[0] + .. + v[n - 1]
hipError_t launch_sum_u32(const uint32_t* v, uint32_t n, unsigned long long* acc, hipStream_t s);
hipError_t launch_update_neighbors(uint32_t* nbrs, uint32_t nbr_stride, uint32_t max_degree,
                                   const uint32_t* points, uint32_t np, const uint32_t* lists,
                                   const uint32_t* counts, uint32_t R, hipStream_t s);
}  // namespace pann
