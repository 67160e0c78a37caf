/*
 * pann/beam_search.h -- the reference's C++ search surface over libpann.
 *
 * Keeps the names, parameter order and QueryParams fields/constructors of
 * ParlayANN's search API (namespace parlayANN):
 *   QueryParams                 types.h:233-268
 *   stats<indexType>            stats.h:121-155
 *   beam_search                 beamSearch.h:419-445 (one query, frontier)
 *   beam_search_rerank          beamSearch.h:528-597 (one query, k results)
 *   qsearchAll                  beamSearch.h:599-646 (a batch)
 * std::vector replaces parlay::sequence (parlaylib is not a dependency).  The
 * point set and graph live on the GPU in a DeviceIndex (the analogue of
 * PointRange + Graph after loading, point_range.h:74-117, graph.h:151-208);
 * every search runs there, and there is no CPU path.  Errors the reference
 * reports with abort() throw std::runtime_error carrying pann_last_error().
 *
 * Header-only; link with libpann.so (parlayann_amd/libpann.so).
 */
#ifndef PANN_BEAM_SEARCH_HPP
#define PANN_BEAM_SEARCH_HPP

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../pann.h"

namespace parlayANN {

/* QueryParams (types.h:233-268): the fields and both constructors.  The
 * early-stopping and range fields are carried but not used by this search. */
enum rangeQueryType { None, Greedy, Doubling, Beam };

struct QueryParams {
  long k;
  long beamSize;
  long limit;
  long degree_limit;
  double rerank_factor = 100;
  double batch_factor = .125;
  bool is_early_stop = false;
  double early_stopping_radius;
  double early_stopping_count;
  rangeQueryType range_query_type = None;
  double radius;
  float pad = 1.0;

  QueryParams(long k, long Q, long limit, long dg, double rerank_factor = 100,
              double batch_factor = .125)
      : k(k), beamSize(Q), limit(limit), degree_limit(dg), rerank_factor(rerank_factor),
        batch_factor(batch_factor) {}

  QueryParams(long k, long Q, long limit, long dg, long es, double esr, long esc,
              rangeQueryType range_query_type, double radius)
      : k(k), beamSize(Q), limit(limit), degree_limit(dg), is_early_stop(es),
        early_stopping_radius(esr), early_stopping_count(esc),
        range_query_type(range_query_type), radius(radius) {}

  QueryParams() {}
};

/* stats<indexType> (stats.h:121-155): per-query visited and distance counts. */
template <typename indexType>
struct stats {
  stats() {}
  explicit stats(size_t n) : visited(n, 0), distances(n, 0) {}

  std::vector<indexType> visited;
  std::vector<indexType> distances;

  void increment_dist(indexType i, indexType j) { distances[i] += j; }
  void increment_visited(indexType i, indexType j) { visited[i] += j; }
  std::vector<indexType> visited_stats() { return statistics(visited); }
  std::vector<indexType> dist_stats() { return statistics(distances); }
  void clear() {
    std::fill(visited.begin(), visited.end(), 0);
    std::fill(distances.begin(), distances.end(), 0);
  }
  /* {average, 99th percentile} (stats.h:147-153) */
  static std::vector<indexType> statistics(std::vector<indexType> s) {
    if (s.empty()) return {0, 0};
    long sum = 0;
    for (auto x : s) sum += (long)x;
    auto avg = (indexType)(sum / (long)s.size());
    std::sort(s.begin(), s.end());
    auto tail = s[(size_t)(.99 * (float)s.size())];
    return {avg, tail};
  }
};

/* Row-major points on the host (the data of a PointRange, without its 64-byte
 * row padding): n rows of dims elements of T (float, uint8_t or int8_t). */
template <typename T>
struct PointRange {
  const T *values = nullptr;
  size_t n = 0;
  unsigned dims = 0;
  PointRange() {}
  PointRange(const T *values, size_t n, unsigned dims) : values(values), n(n), dims(dims) {}
  size_t size() const { return n; }
  unsigned dimension() const { return dims; }
  unsigned get_dims() const { return dims; }
  const T *operator[](size_t i) const { return values + i * dims; }
};

/* Graph<indexType> storage (graph.h:46-64): n rows of maxDeg + 1 entries,
 * slot 0 the degree. */
template <typename indexType>
struct Graph {
  const indexType *adj = nullptr;
  size_t n = 0;
  unsigned maxDeg = 0;
  Graph() {}
  Graph(const indexType *adj, size_t n, unsigned maxDeg) : adj(adj), n(n), maxDeg(maxDeg) {}
  size_t size() const { return n; }
  unsigned max_degree() const { return maxDeg; }
};

enum class Metric { Euclidian, Mips };

namespace detail {
inline void check(pann_status st) {
  if (st != PANN_OK) throw std::runtime_error(std::string("pann: ") + pann_last_error());
}
template <typename T>
constexpr int32_t dtype_of() {
  static_assert(std::is_same<T, float>::value || std::is_same<T, uint8_t>::value ||
                    std::is_same<T, int8_t>::value,
                "points must be float, uint8_t or int8_t");
  return std::is_same<T, float>::value ? PANN_DT_F32
         : std::is_same<T, uint8_t>::value ? PANN_DT_U8 : PANN_DT_I8;
}
inline pann_qparams qparams(const QueryParams &QP) {
  return pann_qparams{QP.k, QP.beamSize, QP.limit, QP.degree_limit, QP.rerank_factor,
                      QP.batch_factor};
}
}  // namespace detail

/* Points and graph resident on one GPU.  Movable, not copyable. */
template <typename T>
class DeviceIndex {
 public:
  DeviceIndex(const PointRange<T> &Points, const Graph<uint32_t> &G, Metric metric,
              int device = 0)
      : n_(Points.size()), dims_(Points.dimension()), maxDeg_(G.max_degree()) {
    if (G.size() != Points.size())
      throw std::runtime_error("graph size and point size do not match");
    pann_index_desc d{(uint32_t)n_, dims_, maxDeg_, detail::dtype_of<T>(),
                      metric == Metric::Euclidian ? PANN_METRIC_L2 : PANN_METRIC_MIPS,
                      Points.values, G.adj, device};
    detail::check(pann_open(&d, &h_));
  }
  DeviceIndex(const DeviceIndex &) = delete;
  DeviceIndex &operator=(const DeviceIndex &) = delete;
  DeviceIndex(DeviceIndex &&o) noexcept : h_(o.h_), n_(o.n_), dims_(o.dims_), maxDeg_(o.maxDeg_) {
    o.h_ = nullptr;
  }
  ~DeviceIndex() {
    if (h_) pann_close(h_);
  }
  /* reduced-precision gather copy of float points: PANN_STORE_BF16 / _FP8 */
  void build_store(int32_t store) { detail::check(pann_build_store(h_, store)); }
  pann_index_t handle() const { return h_; }
  size_t size() const { return n_; }
  unsigned dimension() const { return dims_; }
  unsigned max_degree() const { return maxDeg_; }

 private:
  pann_index_t h_ = nullptr;
  size_t n_;
  unsigned dims_, maxDeg_;
};

/* Search options beyond QueryParams: which store the search gathers, FP32
 * summation order (in-order = the reference's bit pattern), and whether
 * results are re-scored against the float points (the rerank branch of
 * beam_search_rerank, beamSearch.h:569-587). */
struct SearchOptions {
  int32_t store = PANN_STORE_NATIVE;
  int32_t summation = PANN_SUM_INORDER;
  bool rerank = false;
  bool exact_counts = true;  /* QueryStats.distances = the reference's dist_cmps */
};

/* qsearchAll (beamSearch.h:599-646) from one starting point: the k nearest
 * ids per query (beam_search_rerank per query, :528-597), with visited and
 * distance counts in QueryStats when it is sized for the queries. */
template <typename T>
std::vector<std::vector<uint32_t>> qsearchAll(const PointRange<T> &Query_Points,
                                              DeviceIndex<T> &index,
                                              stats<uint32_t> &QueryStats,
                                              const uint32_t starting_point,
                                              const QueryParams &QP,
                                              const SearchOptions &so = SearchOptions()) {
  const size_t nq = Query_Points.size();
  const size_t k = (size_t)std::max<long>(QP.k, 0);
  std::vector<uint32_t> ids(nq * k);
  std::vector<float> dists(nq * k);
  std::vector<uint32_t> vis(nq), cmps(nq);
  pann_qparams qp = detail::qparams(QP);
  pann_search_opts o{so.store, so.summation, so.rerank ? 1 : 0, 0, 0, 0, so.exact_counts ? 1 : 0};
  uint32_t start = starting_point;
  if (nq)
    detail::check(pann_search_host(index.handle(), Query_Points.values, (uint32_t)nq, &qp,
                                   &start, 1, &o, ids.data(), dists.data(), vis.data(),
                                   cmps.data()));
  std::vector<std::vector<uint32_t>> all(nq);
  for (size_t i = 0; i < nq; i++) all[i].assign(ids.begin() + i * k, ids.begin() + (i + 1) * k);
  if (QueryStats.visited.size() == nq) {
    for (size_t i = 0; i < nq; i++) {
      QueryStats.increment_visited((uint32_t)i, vis[i]);
      QueryStats.increment_dist((uint32_t)i, cmps[i]);
    }
  }
  return all;
}

/* beam_search_rerank (beamSearch.h:528-597) for one query: k (id, distance)
 * pairs.  Batch queries through qsearchAll: one call launches one kernel. */
template <typename T>
std::vector<std::pair<uint32_t, float>> beam_search_rerank(
    const T *p, DeviceIndex<T> &index, stats<uint32_t> &QueryStats, uint32_t query_id,
    const std::vector<uint32_t> &starting_points, const QueryParams &QP,
    const SearchOptions &so = SearchOptions()) {
  const size_t k = (size_t)std::max<long>(QP.k, 0);
  std::vector<uint32_t> ids(k);
  std::vector<float> dists(k);
  uint32_t vis = 0, cmps = 0;
  pann_qparams qp = detail::qparams(QP);
  pann_search_opts o{so.store, so.summation, so.rerank ? 1 : 0, 0, 0, 0, so.exact_counts ? 1 : 0};
  detail::check(pann_search_host(index.handle(), p, 1, &qp, starting_points.data(),
                                 (uint32_t)starting_points.size(), &o, ids.data(), dists.data(),
                                 &vis, &cmps));
  if (query_id < QueryStats.visited.size()) {
    QueryStats.increment_visited(query_id, vis);
    QueryStats.increment_dist(query_id, cmps);
  }
  std::vector<std::pair<uint32_t, float>> r(k);
  for (size_t i = 0; i < k; i++) r[i] = {ids[i], dists[i]};
  return r;
}

/* beam_search (beamSearch.h:419-445) for one query: ((frontier, visited),
 * dist_cmps).  The frontier is the final beam, sorted by (distance, id); the
 * visited list is not materialised on this path and comes back empty. */
template <typename T>
std::pair<std::pair<std::vector<std::pair<uint32_t, float>>,
                    std::vector<std::pair<uint32_t, float>>>,
          size_t>
beam_search(const T *p, DeviceIndex<T> &index, const std::vector<uint32_t> &starting_points,
            const QueryParams &QP, const SearchOptions &so = SearchOptions()) {
  const size_t B = (size_t)std::max<long>(QP.beamSize, 0);
  std::vector<uint32_t> ids(B);
  std::vector<float> dists(B);
  uint32_t cnt = 0, vis = 0, cmps = 0;
  pann_qparams qp = detail::qparams(QP);
  pann_search_opts o{so.store, so.summation, 0, 0, 0, 0, so.exact_counts ? 1 : 0};
  detail::check(pann_beam_search_host(index.handle(), p, 1, &qp, starting_points.data(),
                                      (uint32_t)starting_points.size(), &o, ids.data(),
                                      dists.data(), &cnt, &vis, &cmps));
  std::vector<std::pair<uint32_t, float>> frontier(cnt);
  for (uint32_t i = 0; i < cnt; i++) frontier[i] = {ids[i], dists[i]};
  return {{std::move(frontier), {}}, (size_t)cmps};
}

}  // namespace parlayANN

#endif
