// Drives include/pann/beam_search.h the way a reference caller would
// (checkRecall -> qsearchAll, a single beam_search, a beam_search_rerank) and
// writes what it got, for the GPU test to compare with the oracle.
//
// usage: header_check <in> <out>
//   in : u32 n, dim, R, nq, k, B, start; f32 points[n*dim]; u32 adj[n*(R+1)];
//        f32 queries[nq*dim]
//   out: u32 ids[nq*k], visited[nq], cmps[nq]            (qsearchAll)
//        u32 fcount; u32 fids[fcount]; f32 fdists[fcount]; u64 fcmps
//                                                        (beam_search, query 0)
//        u32 rids[k]; f32 rdists[k]                      (beam_search_rerank, query 0)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pann/beam_search.h"

static void rd(FILE* f, void* p, size_t n) {
  if (fread(p, 1, n, f) != n) {
    fprintf(stderr, "short read\n");
    exit(2);
  }
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: header_check <in> <out>\n");
    return 2;
  }
  FILE* in = fopen(argv[1], "rb");
  if (!in) return 2;
  uint32_t h[7];
  rd(in, h, sizeof(h));
  const uint32_t n = h[0], dim = h[1], R = h[2], nq = h[3], k = h[4], B = h[5], start = h[6];
  std::vector<float> pts((size_t)n * dim), qs((size_t)nq * dim);
  std::vector<uint32_t> adj((size_t)n * (R + 1));
  rd(in, pts.data(), pts.size() * 4);
  rd(in, adj.data(), adj.size() * 4);
  rd(in, qs.data(), qs.size() * 4);
  fclose(in);

  using namespace parlayANN;
  PointRange<float> Points(pts.data(), n, dim), Queries(qs.data(), nq, dim);
  Graph<uint32_t> G(adj.data(), n, R);
  DeviceIndex<float> index(Points, G, Metric::Euclidian);
  QueryParams QP(k, B, n, R);  // the CLI's QueryParams(k, Q, n, maxDeg) (check_nn_recall.h:213)
  stats<uint32_t> QueryStats(nq);
  auto all = qsearchAll(Queries, index, QueryStats, start, QP);
  auto bs = beam_search(Queries[0], index, std::vector<uint32_t>{start}, QP);
  stats<uint32_t> one(1);
  auto rr = beam_search_rerank(Queries[0], index, one, 0, std::vector<uint32_t>{start}, QP);

  FILE* out = fopen(argv[2], "wb");
  if (!out) return 2;
  for (auto& v : all) fwrite(v.data(), 4, v.size(), out);
  fwrite(QueryStats.visited.data(), 4, nq, out);
  fwrite(QueryStats.distances.data(), 4, nq, out);
  const auto& fr = bs.first.first;
  uint32_t fc = (uint32_t)fr.size();
  fwrite(&fc, 4, 1, out);
  for (auto& e : fr) fwrite(&e.first, 4, 1, out);
  for (auto& e : fr) fwrite(&e.second, 4, 1, out);
  uint64_t c = bs.second;
  fwrite(&c, 8, 1, out);
  for (auto& e : rr) fwrite(&e.first, 4, 1, out);
  for (auto& e : rr) fwrite(&e.second, 4, 1, out);
  fclose(out);
  return 0;
}
