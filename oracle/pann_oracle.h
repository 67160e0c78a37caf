/*
 * pann_oracle.h -- CPU restatement of ParlayANN's query hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * (or as the timed host-CPU baseline).  The product (parlayann_amd, libpann)
 * never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the ParlayANN source tree).  Arithmetic follows the reference exactly:
 *   - float L2/IP: strictly in-order left-to-right sum; the terms up to the
 *     last multiple of 4 are separately rounded products added one at a
 *     time, the last d % 4 terms are fused multiply-adds -- the codegen of
 *     g++ 11.4 -O3 -march=<FMA target> for the reference's loops (built here
 *     with -ffp-contract=off, the fused terms written as fmaf);
 *   - uint8/int8: int32 accumulate then (float) conversion.
 */
#ifndef PANN_ORACLE_H
#define PANN_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_F32 = 0, OR_U8 = 1, OR_I8 = 2 };
enum { OR_L2 = 0, OR_IP = 1 };

/* Distances (euclidian_point.h:45-90, mips_point.h:44-66). p = base row, q = query. */
float oracle_distance(const void *p, const void *q, unsigned d, int dtype, int metric);

/* hashset<uint32_t> (hashset.h:12-55), exposed for property tests. */
typedef struct oracle_hashset oracle_hashset;
oracle_hashset *oracle_hashset_new(long n);
int oracle_hashset_test(oracle_hashset *h, uint32_t a); /* 1 = already present */
long oracle_hashset_num_entries(const oracle_hashset *h);
long oracle_hashset_capacity(const oracle_hashset *h);
int oracle_hashset_grew(const oracle_hashset *h);
void oracle_hashset_free(oracle_hashset *h);

/* A point store: n rows of `dim` elements of `dtype`, row stride in bytes. */
typedef struct {
  const void *data;
  size_t n;
  unsigned dim;
  size_t stride; /* bytes between rows */
  int dtype;
  int metric;
} oracle_points;

/* Fixed-stride adjacency: n rows of (max_degree + 1) uint32, slot 0 = degree
 * (graph.h:46-64, 130-145). */
typedef struct {
  const uint32_t *adj;
  size_t n;
  unsigned max_degree;
} oracle_graph;

typedef struct {
  long k, beam_size, limit, degree_limit;
  double rerank_factor, batch_factor;
} oracle_qparams;

/* filtered_beam_search without filtering (beamSearch.h:33-228).
 *  - frontier_* receive up to beam_size entries (sorted by (dist,id));
 *  - visited_* receive the sorted visited list if visited_cap allows (the
 *    full count is always returned in *visited_size);
 *  - self_id is the id whose row is `query` (Points[a].same_as(p)), or -1.
 * Returns 0 on success, -1 on invalid arguments (the reference aborts).     */
int oracle_beam_search(const oracle_graph *G, const oracle_points *P,
                       const void *query, long self_id,
                       const uint32_t *starts, int nstarts,
                       const oracle_qparams *QP,
                       uint32_t *frontier_ids, float *frontier_dists, int *frontier_size,
                       uint32_t *visited_ids, float *visited_dists, long *visited_size,
                       long visited_cap, uint64_t *dist_cmps, uint64_t *full_dist_cmps,
                       int *hash_grew);

/* qsearchAll (beamSearch.h:599-646) + beam_search_rerank (:528-597) over a
 * batch of queries, fixed start point, `nthreads` host threads.  When `Q`
 * (search store) differs from `P` (full-precision store) the rerank branch is
 * taken (:569-587); pass Q = NULL for the plain branch (:588-596).
 * out_ids/out_dists: nq x k.  out_visited/out_cmps: per-query stats (may be
 * NULL).  Returns 0, or -1 if k > beam_size or a beam returned < k entries.  */
int oracle_search_batch(const oracle_graph *G, const oracle_points *P,
                        const oracle_points *Q, const void *queries, const void *qqueries,
                        size_t nq, uint32_t start, const oracle_qparams *QP, int nthreads,
                        uint32_t *out_ids, float *out_dists, uint32_t *out_visited,
                        uint64_t *out_cmps, int *out_hash_grew);

/* knn_index::robustPrune (index.h:63-120): cand = (id, dist-to-p) pairs,
 * `add` appends p's current out-neighbours.  Writes <= R ids to out, returns
 * the count.  */
int oracle_robust_prune(const oracle_graph *G, const oracle_points *P, uint32_t p,
                        const uint32_t *cand_ids, const float *cand_dists, int ncand,
                        double alpha, int R, int add, uint32_t *out, long *dist_comps);

/* knn_index::build_index (index.h:150-186) with batch_insert (:188-319) on a
 * mutable adjacency array `adj` of n x (R + 1) uint32 (zeroed by the caller).
 * perms: num_passes x n insertion orders (the reference draws
 * parlay::random_permutation; here the caller supplies it).  Reverse-edge
 * groups keep values in insertion order.  sort_neighbors sorts each list by
 * (distance, id).  Returns 0 or -1.                                         */
int oracle_build_vamana(const oracle_points *P, uint32_t *adj, unsigned R, unsigned L,
                        double alpha, int num_passes, double batch_factor,
                        const uint32_t *perms, int sort_neighbors, int nthreads);

/* FP64 brute force top-k (compute_groundtruth.cpp:22-60, but in double and
 * with (dist, id) ordering).  out_ids/out_dists: nq x k. */
int oracle_groundtruth_f64(const oracle_points *P, const void *queries, size_t nq,
                           int k, int nthreads, uint32_t *out_ids, double *out_dists);

/* HNSW layered graph (HNSW.hpp): adj[l] = n x (max_degree[l] + 1) uint32 in
 * Graph layout for levels l = 0..top (degree 2m at 0, m above; node ids). */
typedef struct {
  size_t n;
  int top;
  uint32_t *const *adj;
  const unsigned *max_degree;
} oracle_hnsw;

/* HNSW::search (HNSW.hpp:1568-1613) per query: ef = 1 descent from the
 * entrance through levels top..1 (search_layer_to, :1545-1566), layer-0 beam
 * search with ef and `limit` (<= 0: n), each level's degree limit its
 * max_degree; the frontier cut after the last entry tied with the k-th
 * distance.  out_ids/out_dists: nq x ef (node ids); out_count: entries kept;
 * out_cmps / out_visited summed over the levels.  Returns 0 or -1. */
int oracle_hnsw_search(const oracle_hnsw *H, const oracle_points *P, const uint32_t *entrance,
                       int nentrance, const void *queries, size_t nq, int k, int ef, long limit,
                       int nthreads, uint32_t *out_ids, float *out_dists, int *out_count,
                       uint64_t *out_cmps, uint32_t *out_visited);

/* HNSW::HNSW (HNSW.hpp:748-807) + insert (:809-996): node i is point i, with
 * the level levels[i] (the reference draws them at random; here they are
 * inputs), batches [b, min(n, ceil(b * batch_base) + 1, b + 0.02 n)), per
 * batch a beam-1 descent, then for each level from the entrance's down: a
 * search with efc from the previous level's results, prune_heuristic to the
 * level's max degree, forward lists, and reverse edges grouped by target in
 * insertion order (appended, or the nearest max_degree by (dist, id)).  H->adj
 * must be zeroed with top >= every level.  Writes the final entrance node.  */
int oracle_build_hnsw(const oracle_hnsw *H, const oracle_points *P, const uint32_t *levels,
                      unsigned efc, float alpha, float batch_base, uint32_t *entrance_out,
                      int nthreads);

/* The reference's 8-bit scalar quantization of float points, as its Python
 * quant path builds it (graph_index.cpp:86-99):
 *  - L2: Euclidian_Point<uint8_t> (euclidian_point.h:98-107, 188-243): slope =
 *    255 / (max - min), offset = round(min * slope) over all values (0
 *    included; all-integer data in [0, 255] keeps slope 1, offset 0), codes
 *    clamp(round(x * slope) - offset, 0, 255) (x truncated when slope 1 and
 *    offset 0);
 *  - MIPS: Quantized_Mips_Point<8, true> (mips_point.h:292-470): bound =
 *    max(q_hi, -q_lo) of the 0.01%-trimmed per-point max / min, scale =
 *    127 / bound, codes round(x * scale) clamped to +-127. */
typedef struct {
  float slope;
  int32_t offset;
  float max_val;
  float scale;
} oracle_q8_params;
int oracle_q8_params_compute(const oracle_points *P, oracle_q8_params *out);
void oracle_q8_translate(const oracle_q8_params *prm, int metric, const float *x, unsigned d,
                         void *codes);

/* The (dist, id) sort used everywhere above, exposed for its property test. */
void oracle_sort_pairs(uint32_t *ids, float *dists, long n);

/* The insert schedules of the two builders, exposed for the pin tests:
 * batch_insert (vamana/index.h:202-234) and HNSW::HNSW (HNSW.hpp:776-792).
 * Each returns the number of batches and writes at most cap of them. */
long oracle_vamana_batches(size_t n, size_t m, size_t *floors, size_t *ceilings, long cap);
long oracle_hnsw_batches(uint32_t n, float batch_base, uint32_t *begins, uint32_t *ends,
                         long cap);

#ifdef __cplusplus
}
#endif
#endif
