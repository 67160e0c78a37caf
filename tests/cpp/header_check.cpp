// Drives include/pann/beam_search.h the way a reference caller would, with the
// reference's signatures: Graph and PointRange read from their files,
// qsearchAll (checkRecall's call, check_nn_recall.h:49-57), one beam_search
// and one beam_search_rerank; then the same over an FP8 copy of the points
// (PointRange<FP8_Point<...>>(Points), the quantizing constructor) with the
// FP32 rerank.  Writes what it got, for the GPU test to compare with the
// oracle.
//
// usage: header_check <data.fbin> <graph> <queries.fbin> <k> <beam> <start> <out>
//   out: u32 ids[nq*k], visited[nq], cmps[nq]                (qsearchAll)
//        u32 fcount; u32 fids[fcount]; f32 fdists[fcount]
//        u32 vcount; u32 vids[vcount]; f32 vdists[vcount]; u64 cmps
//                                                             (beam_search, query 0)
//        u32 rids[k]; f32 rdists[k]                           (beam_search_rerank, query 0)
//        u32 qids[nq*k]                                       (qsearchAll, FP8 + rerank)
//        u32 rrids[k]; f32 rrdists[k]       (beam_search_rerank, FP8 + rerank, query 0)
//        u64 allocations made by a repeated qsearchAll
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pann/beam_search.h"

int main(int argc, char** argv) {
  if (argc != 8) {
    fprintf(stderr, "usage: header_check <data> <graph> <queries> <k> <beam> <start> <out>\n");
    return 2;
  }
  using namespace parlayANN;
  using Point = Euclidian_Point<float>;
  using QPoint = FP8_Point<Point>;
  const long k = atol(argv[4]), B = atol(argv[5]);
  const unsigned start = (unsigned)atol(argv[6]);
  PointRange<Point> Points(argv[1]);
  Graph<unsigned> G(argv[2]);
  PointRange<Point> Query_Points(argv[3]);
  const size_t nq = Query_Points.size();
  search_options().exact_counts = true;  // QueryStats.distances = the reference's dist_cmps

  // the CLI's QueryParams(k, Q, n, maxDeg) (check_nn_recall.h:213)
  QueryParams QP(k, B, (long)G.size(), G.max_degree());
  stats<unsigned> QueryStats(nq);
  auto all = qsearchAll(Query_Points, Query_Points, Query_Points, G, Points, Points, Points,
                        QueryStats, start, QP);
  auto bs = beam_search(Query_Points[0], G, Points, start, QP);
  stats<unsigned> one(1);
  auto rr = beam_search_rerank(Query_Points[0], Query_Points[0], Query_Points[0], G, Points,
                               Points, Points, one, std::vector<unsigned>{start}, QP);

  search_options().exact_counts = false;
  PointRange<QPoint> Q_Points(Points), Q_Query(Query_Points);
  stats<unsigned> qs(nq);
  auto qall = qsearchAll(Query_Points, Q_Query, Q_Query, G, Points, Q_Points, Q_Points, qs, start,
                         QP);
  auto qrr = beam_search_rerank(Query_Points[0], Q_Query[0], Q_Query[0], G, Points, Q_Points,
                                Q_Points, qs, std::vector<unsigned>{start}, QP, false);
  uint64_t a0 = 0, a1 = 0;
  pann_device_alloc_count(&a0);
  stats<unsigned> qs2(nq);
  auto again = qsearchAll(Query_Points, Q_Query, Q_Query, G, Points, Q_Points, Q_Points, qs2,
                          start, QP);
  pann_device_alloc_count(&a1);
  if (again != qall) {
    fprintf(stderr, "a repeated qsearchAll returned other ids\n");
    return 3;
  }
  // two waves per query: a launch option only, the same ids
  search_options().waves_per_query = 2;
  stats<unsigned> qs3(nq);
  auto coop = qsearchAll(Query_Points, Q_Query, Q_Query, G, Points, Q_Points, Q_Points, qs3,
                         start, QP);
  search_options().waves_per_query = 1;
  if (coop != qall) {
    fprintf(stderr, "qsearchAll with two waves per query returned other ids\n");
    return 4;
  }

  FILE* out = fopen(argv[7], "wb");
  if (!out) return 2;
  for (auto& v : all) fwrite(v.data(), 4, v.size(), out);
  fwrite(QueryStats.visited.data(), 4, nq, out);
  fwrite(QueryStats.distances.data(), 4, nq, out);
  const auto& fr = bs.first.first;
  const auto& vi = bs.first.second;
  uint32_t fc = (uint32_t)fr.size(), vc = (uint32_t)vi.size();
  fwrite(&fc, 4, 1, out);
  for (auto& e : fr) fwrite(&e.first, 4, 1, out);
  for (auto& e : fr) fwrite(&e.second, 4, 1, out);
  fwrite(&vc, 4, 1, out);
  for (auto& e : vi) fwrite(&e.first, 4, 1, out);
  for (auto& e : vi) fwrite(&e.second, 4, 1, out);
  uint64_t c = (uint64_t)bs.second;
  fwrite(&c, 8, 1, out);
  for (auto& e : rr) fwrite(&e.first, 4, 1, out);
  for (auto& e : rr) fwrite(&e.second, 4, 1, out);
  for (auto& v : qall) fwrite(v.data(), 4, v.size(), out);
  for (auto& e : qrr) fwrite(&e.first, 4, 1, out);
  for (auto& e : qrr) fwrite(&e.second, 4, 1, out);
  uint64_t da = a1 - a0;
  fwrite(&da, 8, 1, out);
  fclose(out);
  return 0;
}
