/*
 * pann.h -- C ABI of the MI355X beam-search library (libpann.so).
 *
 * Plain C types only.  Each entry point names the reference interface it
 * replaces (paths relative to the ParlayANN source tree).  Errors are status
 * codes (the reference prints and calls abort(), e.g. beamSearch.h:51-57,
 * :612-616, graph.h:60-63); pann_last_error() gives the message.
 *
 * Device pointers: arguments documented "device" must be HIP device memory of
 * the index's device; pointers documented "host" are ordinary host memory.
 * `stream` is a hipStream_t (NULL = default stream).
 */
#ifndef PANN_H
#define PANN_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PANN_OK = 0,
  PANN_ERR_ARG = 1,         /* invalid argument (the reference aborts) */
  PANN_ERR_DEVICE = 2,      /* HIP runtime error */
  PANN_ERR_OOM = 3,         /* device allocation failed */
  PANN_ERR_CAPACITY = 4,    /* a query outgrew a fixed capacity (visited list; the exact seen
                               set's table in an async search) */
  PANN_ERR_SHORT_BEAM = 5,  /* beam returned fewer than k entries (beamSearch.h:557-562) */
  PANN_ERR_UNSUPPORTED = 6  /* shape or option this build does not implement */
} pann_status;

/* element type of the stored points (point_range.h:43-141 template argument) */
typedef enum { PANN_DT_F32 = 0, PANN_DT_U8 = 1, PANN_DT_I8 = 2 } pann_dtype;

/* Euclidian_Point (euclidian_point.h:92-248) / Mips_Point (mips_point.h:68-144) */
typedef enum { PANN_METRIC_L2 = 0, PANN_METRIC_MIPS = 1 } pann_metric;

/* which copy of the points the beam search gathers */
typedef enum {
  PANN_STORE_NATIVE = 0, /* the points as given (f32 / u8 / i8) */
  PANN_STORE_BF16 = 1,   /* f32 points rounded to bfloat16 (RNE) */
  PANN_STORE_FP8 = 2,    /* f32 points as OCP FP8 e4m3 with per-dimension absmax scales;
                            inner product sums exact products of the codes and the query
                            term q_j s_j rounded to FP16 (scaled per query by a power of
                            two) in FP32 */
  PANN_STORE_Q8 = 3,     /* the reference's 8-bit scalar quantization of f32 points: uint8
                            affine codes for L2 (Euclidian_Point<uint8_t>, euclidian_point.h:
                            98-107, 188-243), int8 codes with a trimmed global scale for MIPS
                            (Quantized_Mips_Point<8, true>, mips_point.h:292-470); queries are
                            encoded the same way, distances are exact int32 */
  PANN_STORE_FP8_RESIDUAL = 4 /* MIPS only, dim <= 256: x - c as OCP FP8 e4m3 with
                            per-dimension absmax scales, c the nearest (L2) of ncent points
                            sampled from the index; the search adds q.c from a per-query
                            table.  For clustered data, whose neighbour gaps are below
                            e4m3's error on x itself.  No counterpart in the reference,
                            which quantizes x directly (mips_point.h:292-470) */
} pann_store;

/* FP32 summation order of the native f32 store */
typedef enum {
  PANN_SUM_INORDER = 0, /* in-order sum, the reference's float loop (euclidian_point.h:83-90,
                           mips_point.h:60-66) as g++ -O3 -march=<FMA target> compiles it:
                           separately rounded terms up to the last multiple of 4, the last
                           d % 4 terms fused; bit-identical to that build at every d.  Rows
                           are gathered 8 lanes per row and the running sum is carried from
                           lane to lane in element order (DPP) */
  PANN_SUM_TREE = 1     /* 8 lanes per row, DPP tree reduction (faster, not bit-identical) */
} pann_summation;

typedef struct pann_index_s *pann_index_t;

/* An index over n points of `dim` elements and a fixed-degree graph.
 * graph: n x (max_degree + 1) uint32, slot 0 = degree, then the neighbours
 * (Graph<unsigned>, graph.h:130-255; file format graph.h:151-237). */
typedef struct {
  uint32_t n, dim, max_degree;
  int32_t dtype;          /* pann_dtype */
  int32_t metric;         /* pann_metric */
  const void *points;     /* host, n x dim row-major */
  const uint32_t *graph;  /* host, n x (max_degree + 1) */
  int32_t device;         /* HIP device ordinal */
} pann_index_desc;

/* QueryParams (types.h:233-268): k, beamSize, limit, degree_limit,
 * rerank_factor, batch_factor.  Same field meaning and defaults
 * (rerank_factor 100, batch_factor 0.125 -- types.h:237-238). */
typedef struct {
  int64_t k;
  int64_t beam_size;
  int64_t limit;
  int64_t degree_limit;
  double rerank_factor;
  double batch_factor;
} pann_qparams;

typedef struct {
  int32_t store;       /* pann_store gathered by the search */
  int32_t summation;   /* pann_summation for the native f32 store */
  int32_t rerank;      /* 1: rerank branch of beam_search_rerank (beamSearch.h:569-587)
                          against the f32 points; 0: plain branch (:588-596) */
  int32_t hash_slots_log2; /* exact_counts = 1 only: LDS seen-set size, log2 of slots in
                              [8, 15] (0 = automatic) */
  int32_t async;       /* 1: do not synchronise; call pann_search_status() later */
  int32_t waves_per_cu;/* persistent waves per CU (0 = automatic) */
  int32_t exact_counts;/* 1: out_cmps equals the reference's dist_cmps (every id seen is
                          remembered in an LDS table with a per-wave global overflow table,
                          each filled to half its slots; a synchronous search that outgrows
                          the global table grows it 4x and runs again, as the reference's
                          hashset grows past half load, hashset.h:25-38, and the index
                          keeps the larger table; an async search reports
                          PANN_ERR_CAPACITY instead); 0: every neighbour is evaluated and checked
                          against the frontier and pending candidates afterwards -- same
                          ids, distances and visited counts, out_cmps counts evaluations, no
                          capacity limit */
  int32_t visits_per_round; /* 2: while a merge is deferred, the next visit's rows are
                               gathered with the current one's; 0 or 3: the same, but only
                               where the current visit is predicted to defer its merge (rows
                               gathered for a visit that a merge cancels are discarded);
                               1: one visit per round.  Results do not depend on it (a
                               tuning knob) */
  int32_t eval_cache_log2;  /* exact_counts = 0: log2 of the slots of an LDS cache of evaluated
                               ids (0 = none, 6..13); a cached id is not evaluated again.
                               Results do not depend on it (a tuning knob) */
  int32_t frontier_filter;  /* exact_counts = 0: a Bloom filter of the frontier's ids, each hit
                               checked exactly, keeps ids already in the frontier from being
                               gathered again.  0 = automatic (on for float32 rows wider than
                               one 128-byte line), 1 = on, 2 = off.  Results do not depend
                               on it (a tuning knob) */
  int32_t waves_per_query;  /* 0 or 1: one 64-lane wave per query; 2 or 4: a workgroup of that
                               many waves per query, which share each visit's row gathers,
                               the merges and the rerank (for large beams, where one query's
                               work is long and few queries fit a CU).  waves_per_cu then
                               counts waves.  Results do not depend on it (a tuning knob) */
  int32_t query_order;      /* 0: a batch with per-query beams runs longest-first (a device
                               counting sort by descending beam, so that long searches do not
                               start last and set the batch's end); 1: the queries in the order
                               given.  Results do not depend on it (a tuning knob) */
} pann_search_opts;

/* Creates the device index: uploads the points (rows padded to 128 bytes)
 * and the graph.  Replaces PointRange(char*) + Graph(char*) loading
 * (point_range.h:74-117, graph.h:151-208) for the device path. */
pann_status pann_open(const pann_index_desc *desc, pann_index_t *out);

/* Builds a reduced-precision copy of f32 points on the device
 * (PANN_STORE_BF16 or PANN_STORE_FP8).  The reference's analog is the
 * quantizing PointRange constructor (point_range.h:55-66 with
 * Quantized_Mips_Point / Euclidian_Point<uint8_t> parameters). */
pann_status pann_build_store(pann_index_t idx, int32_t store);

/* Copies the FP8 per-dimension scales (dim floats) to host memory. */
pann_status pann_get_fp8_scales(pann_index_t idx, float *scales_host);

/* Builds PANN_STORE_FP8_RESIDUAL with ncent centroids: the points
 * floor(n / ncent) * i, i < ncent; each point's centroid is its nearest in L2
 * (FP32 distances on the brute-force kernel); scales s_j = max_i |x_ij -
 * c_ij| / 448.  ncent: a multiple of 4 in [4, min(4096, n)], 0 = min(2048,
 * n / 64 rounded down to a multiple of 4), at least 4.  n >= 4.  pann_build_store(idx,
 * PANN_STORE_FP8_RESIDUAL) is this with ncent = 0.  Same role as
 * pann_build_store (point_range.h:55-66); no counterpart in the reference. */
pann_status pann_build_fp8_residual(pann_index_t idx, uint32_t ncent);

/* The FP8 residual store's parameters, to host memory (any output but ncent
 * may be NULL): ncent, the centroids (ncent x dim floats), each point's
 * centroid (n) and the per-dimension scales (dim). */
pann_status pann_get_fp8_residual(pann_index_t idx, uint32_t *ncent, float *centroids,
                                  uint32_t *assign, float *scales);

/* Batched beam search, device buffers.  Replaces qsearchAll
 * (beamSearch.h:599-646) over beam_search_rerank (:528-597) over
 * filtered_beam_search (:33-228): one search per query from `starts`.
 *  queries:   device, nq rows of dim elements in the index dtype
 *             (f32 queries for BF16/FP8 stores)
 *  k_pq, beam_pq: optional device int32[nq] per-query k / beamSize (NULL =
 *             uniform qp->k / qp->beam_size)
 *  starts:    host, nstarts unique ids (1 <= nstarts <= min(beam, 64))
 *  out_ids, out_dists: device, nq x out_stride (first k used per row)
 *  out_visited, out_cmps: optional device uint32[nq]: visited count and
 *             distance comparisons (stats<>, stats.h:121-155)            */
pann_status pann_search(pann_index_t idx, const void *queries, uint32_t nq,
                        const pann_qparams *qp, const int32_t *k_pq, const int32_t *beam_pq,
                        const uint32_t *starts, uint32_t nstarts, const pann_search_opts *opts,
                        uint32_t *out_ids, float *out_dists, uint32_t out_stride,
                        uint32_t *out_visited, uint32_t *out_cmps, void *stream);

/* pann_search without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as pann_search; the memory the launch writes besides the outputs
 * (work counter, status word, start points, PANN_STORE_Q8 query codes) is a
 * file-scope device array of 8 MiB per GPU, held until the process ends, so
 * the index's stream contexts are neither created nor used.  Limits:
 * synchronous only (opts->async = 0); exact_counts = 0; no
 * PANN_STORE_FP8_RESIDUAL; PANN_STORE_Q8 needs nq * dim <= 8 MiB - 1 KiB
 * (PANN_ERR_CAPACITY beyond).  Calls share each GPU's array: this library
 * serialises the calls of one process, and calls that overlap on one GPU in
 * any other way race. */
pann_status pann_search_noalloc(pann_index_t idx, const void *queries, uint32_t nq,
                                const pann_qparams *qp, const int32_t *k_pq,
                                const int32_t *beam_pq, const uint32_t *starts,
                                uint32_t nstarts, const pann_search_opts *opts,
                                uint32_t *out_ids, float *out_dists, uint32_t out_stride,
                                uint32_t *out_visited, uint32_t *out_cmps, void *stream);

/* Same search with host buffers (copies in and out, synchronises).  This is
 * what a C++ qsearchAll caller binds directly.  The device staging for the
 * copies is kept per index and grown geometrically, so repeated calls of the
 * same or smaller size allocate nothing.  Calls on one index are serialised. */
pann_status pann_search_host(pann_index_t idx, const void *queries, uint32_t nq,
                             const pann_qparams *qp, const uint32_t *starts, uint32_t nstarts,
                             const pann_search_opts *opts, uint32_t *out_ids, float *out_dists,
                             uint32_t *out_visited, uint32_t *out_cmps);

/* Status of the last asynchronous pann_search on `stream` (synchronises). */
pann_status pann_search_status(pann_index_t idx, void *stream);

/* filtered_beam_search without filtering (beamSearch.h:33-228) for a batch
 * of host queries from one start set, returning each query's whole final
 * frontier: out_ids/out_dists nq x beam_size (host), out_count[q] entries,
 * sorted by (dist, id) -- the first member of beam_search's result
 * (beamSearch.h:419-445).  A frontier shorter than beam_size is a result.
 * opts.rerank is ignored.  out_visited / out_cmps may be NULL. */
pann_status pann_beam_search_host(pann_index_t idx, const void *queries, uint32_t nq,
                                  const pann_qparams *qp, const uint32_t *starts,
                                  uint32_t nstarts, const pann_search_opts *opts,
                                  uint32_t *out_ids, float *out_dists, uint32_t *out_count,
                                  uint32_t *out_visited, uint32_t *out_cmps);

/* pann_beam_search_host plus each query's visited list -- the second member
 * of filtered_beam_search's result (beamSearch.h:225-227): out_vis_ids /
 * out_vis_dists nq x vis_cap (host), out_vis_count[q] entries, sorted by
 * (dist, id).  vis_cap in [1, 4096]; a longer list is PANN_ERR_CAPACITY. */
pann_status pann_beam_search_visited_host(pann_index_t idx, const void *queries, uint32_t nq,
                                          const pann_qparams *qp, const uint32_t *starts,
                                          uint32_t nstarts, const pann_search_opts *opts,
                                          uint32_t *out_ids, float *out_dists,
                                          uint32_t *out_count, uint32_t vis_cap,
                                          uint32_t *out_vis_ids, float *out_vis_dists,
                                          uint32_t *out_vis_count, uint32_t *out_cmps);

/* Distance evaluations made by this index's build steps since it was created
 * (or since the last call with reset != 0): *search_evals by the build-mode
 * searches (pann_search_visited), *prune_evals by robustPrune
 * (pann_robust_prune_batch / pann_robust_prune_groups, each row distance it
 * computes).  Either pointer may be NULL.  Waits for the device.  No
 * reference counterpart: times rows per evaluation give the build's
 * algorithmic bytes. */
pann_status pann_build_eval_counts(pann_index_t idx, uint64_t *search_evals,
                                   uint64_t *prune_evals, int32_t reset);

/* Number of device allocations this library has made since it was loaded
 * (every hipMalloc of the index, its stores and its per-stream scratch).  No
 * reference counterpart: it lets a caller check that steady-state searches
 * allocate nothing. */
pann_status pann_device_alloc_count(uint64_t *count);

/* Duration in milliseconds of the last pann_search kernel, measured with
 * HIP events on the launch stream (synchronises). */
pann_status pann_last_kernel_ms(pann_index_t idx, float *ms);

/* Build-mode search of Vamana batch insertion: QueryParams(0, L, n,
 * max_degree, ..., batch_factor) from start point `start` for each id in
 * self_ids (the query row is the point itself, same_as skips it), writing
 * the visited list sorted by (dist, id) -- the candidate set robustPrune
 * consumes (vamana/index.h:247-267).
 *  self_ids: device uint32[nq]; vis_ids/vis_dists: device nq x vis_cap;
 *  vis_count: device uint32[nq]                                          */
pann_status pann_search_visited(pann_index_t idx, const uint32_t *self_ids, uint32_t nq,
                                uint32_t beam, double batch_factor, uint32_t start,
                                uint32_t vis_cap, uint32_t *vis_ids, float *vis_dists,
                                uint32_t *vis_count, void *stream);

/* knn_index::robustPrune (vamana/index.h:63-137) for a batch of points on
 * the device: for point p = pts[i], the candidates cand_ids[i*cand_stride
 * .. + cand_count[i]) with distances cand_dists (to p; NULL = computed, the
 * id-only overload :124-137).  add = 0: candidates only; 1: plus p's current
 * out-neighbours (:72-77); 2: the reverse-edge rule of batch_insert
 * (:292-303): if count + degree <= R the candidates followed by the current
 * neighbours not among them (add_neighbors_without_repeats, :140-146), else
 * robustPrune with add; 3: no pruning, p's current out-neighbours sorted by
 * (dist, id) (the final G[i].sort of build_index, :180-185; cand_count must
 * be 0).  Writes <= R ids to out_ids[i*R ..] and the count to out_count[i].
 * Points must be distinct.  All pointers device.                          */
pann_status pann_robust_prune_batch(pann_index_t idx, const uint32_t *pts, uint32_t np,
                                    const uint32_t *cand_ids, const float *cand_dists,
                                    const uint32_t *cand_count, uint32_t cand_stride,
                                    double alpha, uint32_t R, int32_t add, uint32_t *out_ids,
                                    uint32_t *out_count, void *stream);

/* The same robustPrune for groups in CSR form: point pts[i] takes the
 * id-only candidates cand_ids[offsets[i] .. offsets[i+1]) (distances are
 * computed).  This is the shape of batch_insert's reverse-edge step, whose
 * groups come from group_by_key (vamana/index.h:281-303); max_count bounds
 * every group's size.  All pointers device.                               */
pann_status pann_robust_prune_groups(pann_index_t idx, const uint32_t *pts, uint32_t np,
                                     const uint32_t *offsets, const uint32_t *cand_ids,
                                     uint32_t max_count, double alpha, uint32_t R, int32_t add,
                                     uint32_t *out_ids, uint32_t *out_count, void *stream);

/* Replaces the out-neighbour lists of points pts[0..np) on the device
 * (edgeRange::update_neighbors, graph.h:77-88).  lists: device np x R,
 * counts: device uint32[np]. */
pann_status pann_update_neighbors(pann_index_t idx, const uint32_t *pts, uint32_t np,
                                  const uint32_t *lists, const uint32_t *counts, uint32_t R,
                                  void *stream);

/* Parameters of the PANN_STORE_Q8 copy (after pann_build_store): L2 slope
 * and offset, MIPS bound (max_val) and scale; unused fields are 0. */
pann_status pann_get_q8_params(pann_index_t idx, float *slope, int32_t *offset, float *max_val,
                               float *scale);

/* Brute-force candidates for groundtruth (compute_groundtruth.cpp:22-60) on
 * the matrix cores (v_mfma_f32_32x32x2f32): for each of nq device queries
 * (float, dim values each) the kk points of smallest FP32 distance, sorted,
 * as device nq x kk ids (0xFFFFFFFF past n) and distances (L2: |x|^2 - 2 q.x,
 * i.e. without the query's |q|^2; MIPS: -q.x).  kk <= 224.  The exact FP64
 * re-scoring of the candidates, with kk - k of margin for FP32 rounding, is
 * the caller's.  dim <= 256 (PANN_ERR_UNSUPPORTED beyond).  Synchronises
 * `stream`. */
pann_status pann_bruteforce_candidates(pann_index_t idx, const float *queries, uint32_t nq,
                                      uint32_t kk, uint32_t *out_ids, float *out_dists,
                                      void *stream);

/* ------------------------------------------------------------------------
 * HNSW (algorithms/HNSW/HNSW.hpp).  A layered index is a pann_index whose
 * level-0 graph (pann_open, degree 2m) is joined by one graph per upper level
 * (degree m) over the same node ids; node ids are row indices of the points.
 * ------------------------------------------------------------------------ */

/* Sets the graph of level `level` >= 1: n x (max_degree + 1) rows in Graph
 * layout (degree, then node ids; NULL = every list empty).  Replaces the
 * per-level neighbour lists the model loader reads (HNSW.hpp:715-730). */
pann_status pann_hnsw_set_level(pann_index_t idx, uint32_t level, uint32_t max_degree,
                                const uint32_t *graph_host);

/* HNSW::search_layer (HNSW.hpp:1088-1109) for a batch: beam search on graph
 * `level` with the intended QueryParams(ef, ef, limit, degree_limit) (the
 * reference passes a literal 1.35 that lands in `limit`).  Query q starts
 * from starts[q * starts_stride .. + nstarts[q]) (nstarts NULL: one start),
 * 1 <= count <= ef.  queries: device rows; or NULL with self_ids: query q is
 * point self_ids[q], which same_as skips (the insert step, :896-910).  Writes
 * the first min(out_stride, |frontier|) entries, sorted by (dist, id), and
 * |frontier| to out_count.  limit <= 0 means n.  All pointers device. */
pann_status pann_hnsw_search_layer(pann_index_t idx, uint32_t level, const void *queries,
                                   const uint32_t *self_ids, uint32_t nq, uint32_t ef,
                                   uint32_t degree_limit, int64_t limit, const uint32_t *starts,
                                   const uint32_t *nstarts, uint32_t starts_stride,
                                   const pann_search_opts *opts, uint32_t *out_ids,
                                   float *out_dists, uint32_t out_stride, uint32_t *out_count,
                                   uint32_t *out_visited, uint32_t *out_cmps, void *stream);

/* HNSW::search (HNSW.hpp:1568-1613) with search_layer_to (:1545-1566) for a
 * batch of device queries: beam 1 from the entrance (host array, <= 64 ids)
 * through levels top_level .. 1, keeping the closest, then layer 0 with beam
 * ef and `limit` (<= 0: n); each level's degree limit is its max_degree.  The
 * frontier is sorted by distance and, when longer than k, cut after the last
 * entry tied with the k-th distance (:1590-1596).  Ids are mapped through
 * node_ids (device, NULL = node ids).  out_stride >= ef; out_count[q] =
 * entries written; out_visited / out_cmps sum every level (count_cmps).
 * opts.rerank re-scores layer 0 in FP32 before the cut.  k_pq / ef_pq
 * (device int32[nq], may be NULL) give each query its own k and ef <= ef,
 * so a batch mixes requests. */
pann_status pann_hnsw_search(pann_index_t idx, const void *queries, uint32_t nq, uint32_t k,
                             uint32_t ef, const int32_t *k_pq, const int32_t *ef_pq,
                             int64_t limit, const uint32_t *entrance,
                             uint32_t nentrance, uint32_t top_level, const uint32_t *node_ids,
                             const pann_search_opts *opts, uint32_t *out_ids, float *out_dists,
                             uint32_t out_stride, uint32_t *out_count, uint32_t *out_visited,
                             uint32_t *out_cmps, void *stream);

/* select_neighbors / prune_heuristic (HNSW.hpp:430-503) for a batch: point
 * pts[i] takes the candidates cand_ids/cand_dists[i*cand_stride .. +
 * cand_count[i]) (a search frontier with distances to the point), sorted by
 * distance; c is dropped when an already selected r has d(c, r) < alpha *
 * d(c, p) (float compare); at most M are kept.  All pointers device. */
pann_status pann_hnsw_select_neighbors(pann_index_t idx, uint32_t level, const uint32_t *pts,
                                       uint32_t np, const uint32_t *cand_ids,
                                       const float *cand_dists, const uint32_t *cand_count,
                                       uint32_t cand_stride, float alpha, uint32_t M,
                                       uint32_t *out_ids, uint32_t *out_count, void *stream);

/* The reverse-edge step of HNSW::insert (HNSW.hpp:935-983): point pts[i]
 * receives the ids cand_ids[offsets[i] .. offsets[i+1]) on graph `level`;
 * when its list and the new ids fit in M they are appended in order,
 * otherwise both are sorted by distance to the point and cut to M.  Writes
 * the new lists (use pann_hnsw_set_neighbors to apply).  All pointers device. */
pann_status pann_hnsw_add_reverse(pann_index_t idx, uint32_t level, const uint32_t *pts,
                                  uint32_t np, const uint32_t *offsets, const uint32_t *cand_ids,
                                  uint32_t max_count, uint32_t M, uint32_t *out_ids,
                                  uint32_t *out_count, void *stream);

/* neighbourhood(u, level) = list (HNSW.hpp:928-933): lists np x R, counts np. */
pann_status pann_hnsw_set_neighbors(pann_index_t idx, uint32_t level, const uint32_t *pts,
                                    uint32_t np, const uint32_t *lists, const uint32_t *counts,
                                    uint32_t R, void *stream);

/* Copies graph `level` back as n x (max_degree + 1) (for HNSW::save, :1615-1686). */
pann_status pann_hnsw_download_level(pann_index_t idx, uint32_t level, uint32_t *graph_host);

/* Copies the device graph back to host memory as n x (max_degree + 1). */
pann_status pann_download_graph(pann_index_t idx, uint32_t *graph_host);

pann_status pann_close(pann_index_t idx);

/* Message of the last error on this thread. */
const char *pann_last_error(void);

/* Device properties the library sizes its launches by. */
pann_status pann_device_info(int32_t device, int32_t *num_cus, int64_t *lds_per_cu,
                             int64_t *hbm_bytes);

#ifdef __cplusplus
}
#endif
#endif /* PANN_H */
