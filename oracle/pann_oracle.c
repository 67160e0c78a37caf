/*
 * pann_oracle.c -- CPU restatement of ParlayANN's query hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pann_oracle.h).  Built with
 * -ffp-contract=off, so every float operation below rounds where it is
 * written: the float distances are the arithmetic that g++ 11.4 -O3
 * -march=<any FMA target> makes of the reference's loops
 * (euclidian_point.h:83-90, mips_point.h:60-66).  That compiler keeps the
 * loop's order (no -fassociative-math): an 8-wide then a 4-wide vector body
 * forms separately rounded products and adds them one at a time, and the
 * last d % 4 terms run as scalar code whose multiply and add it contracts
 * into one fused multiply-add.  tests/test_f32_reference_loop.py compiles
 * the reference loop with those flags and checks this model at every d.
 */
#define _GNU_SOURCE
#include "pann_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ORACLE_MUTANT plants one plausible mistake in the graph builders, so that
 * the builder pin tests can show each one is caught (tests/test_builder_pins.py
 * builds every variant with `make mutant`).  0, the default, is the oracle.
 *   1  Vamana reverse edges: append only while newsize < R (index.h:295 has <=)
 *   2  robustPrune: prune only when alpha * d(p*, p') < d(p, p') (index.h:111 has <=)
 *   3  Vamana: every point in one batch (index.h:223-234 doubles the batches)
 *   4  Vamana: batches of one point
 *   5  robustPrune ignores p's current out-neighbours (index.h:72-77 adds them)
 *   6  Vamana reverse edges: the current neighbours first, then the new ones
 *      (add_neighbors_without_repeats appends to the new ones, index.h:140-146)
 *   7  HNSW prune_heuristic: prune when d(c, r) <= alpha * d(c, u) (HNSW.hpp:473 has <)
 *   8  HNSW reverse edges: cut when the total reaches m (HNSW.hpp:966 has >)
 *   9  HNSW reverse edges: the cut keeps the farthest m (HNSW.hpp:975 keeps the nearest)
 *   10 HNSW batches end at ceil(begin * base), without the + 1 (HNSW.hpp:781) */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

/* ------------------------------------------------------------------------ */
/* distances: euclidian_point.h:45-90, mips_point.h:44-66                    */
/* ------------------------------------------------------------------------ */

/* euclidian_point.h:83-90: result += (q[i]-p[i])*(q[i]-p[i]), in order;
 * separately rounded terms up to the last multiple of 4, fused ones after. */
static float l2_f32(const float *p, const float *q, unsigned d) {
  float r = 0.0f;
  const unsigned body = d - d % 4;
  for (unsigned i = 0; i < body; i++) {
    float t = q[i] - p[i];
    float s = t * t;
    r = r + s;
  }
  for (unsigned i = body; i < d; i++) {
    float t = q[i] - p[i];
    r = fmaf(t, t, r);
  }
  return r;
}

/* euclidian_point.h:54-62: int32 accumulate of squared differences. */
static float l2_u8(const uint8_t *p, const uint8_t *q, unsigned d) {
  int32_t r = 0;
  for (unsigned i = 0; i < d; i++) {
    int32_t a = (int32_t)p[i], b = (int32_t)q[i];
    r += (a - b) * (a - b);
  }
  return (float)r;
}

/* euclidian_point.h:74-81 */
static float l2_i8(const int8_t *p, const int8_t *q, unsigned d) {
  int32_t r = 0;
  for (unsigned i = 0; i < d; i++) {
    int32_t t = (int32_t)((int16_t)q[i] - (int16_t)p[i]);
    r += t * t;
  }
  return (float)r;
}

/* mips_point.h:60-66: result += q[i]*p[i] in order; return -result.  The
 * last d % 4 terms fused, as for l2_f32. */
static float ip_f32(const float *p, const float *q, unsigned d) {
  float r = 0.0f;
  const unsigned body = d - d % 4;
  for (unsigned i = 0; i < body; i++) {
    float s = q[i] * p[i];
    r = r + s;
  }
  for (unsigned i = body; i < d; i++) r = fmaf(q[i], p[i], r);
  return -r;
}

/* mips_point.h:44-50 */
static float ip_u8(const uint8_t *p, const uint8_t *q, unsigned d) {
  int32_t r = 0;
  for (unsigned i = 0; i < d; i++) r += ((int32_t)q[i]) * ((int32_t)p[i]);
  return -((float)r);
}

/* mips_point.h:52-58 */
static float ip_i8(const int8_t *p, const int8_t *q, unsigned d) {
  int32_t r = 0;
  for (unsigned i = 0; i < d; i++) r += ((int32_t)q[i]) * ((int32_t)p[i]);
  return -((float)r);
}

float oracle_distance(const void *p, const void *q, unsigned d, int dtype, int metric) {
  if (metric == OR_L2) {
    if (dtype == OR_F32) return l2_f32((const float *)p, (const float *)q, d);
    if (dtype == OR_U8) return l2_u8((const uint8_t *)p, (const uint8_t *)q, d);
    return l2_i8((const int8_t *)p, (const int8_t *)q, d);
  }
  if (dtype == OR_F32) return ip_f32((const float *)p, (const float *)q, d);
  if (dtype == OR_U8) return ip_u8((const uint8_t *)p, (const uint8_t *)q, d);
  return ip_i8((const int8_t *)p, (const int8_t *)q, d);
}

static inline const void *row_of(const oracle_points *P, size_t i) {
  return (const char *)P->data + i * P->stride;
}

static inline float pdist(const oracle_points *P, size_t a, const void *q) {
  return oracle_distance(row_of(P, a), q, P->dim, P->dtype, P->metric);
}

/* ------------------------------------------------------------------------ */
/* hashset<uint32_t>: hashset.h:12-55, including its growth quirks           */
/* (probe start keeps the pre-growth mask; num_entries is not reset).        */
/* ------------------------------------------------------------------------ */

struct oracle_hashset {
  int bits;
  uint32_t *entries;
  size_t size;
  size_t mask;
  long num_entries;
  int grew;
};

static const uint32_t HS_EMPTY = 0xFFFFFFFFu;

static void hs_init(struct oracle_hashset *h, long n) {
  h->bits = (int)ceil(log2((double)n)); /* hashset.h:51 */
  h->size = (size_t)1 << h->bits;
  h->entries = (uint32_t *)malloc(h->size * sizeof(uint32_t));
  memset(h->entries, 0xFF, h->size * sizeof(uint32_t));
  h->mask = h->size - 1;
  h->num_entries = 0;
  h->grew = 0;
}

static inline size_t hs_hash(uint32_t k) { return (size_t)((uint64_t)k * UINT64_C(0xbf58476d1ce4e5b9)); }

/* hashset::operator() (hashset.h:22-48) */
static int hs_test(struct oracle_hashset *h, uint32_t a) {
  int loc = (int)(hs_hash(a) & h->mask);
  if (h->entries[loc] == a) return 1;
  if ((size_t)h->num_entries > h->size / 2) { /* :25-38 */
    h->grew = 1;
    h->bits = h->bits + 1;
    size_t nsize = (size_t)1 << h->bits;
    uint32_t *old = h->entries;
    size_t osize = h->size;
    h->entries = (uint32_t *)malloc(nsize * sizeof(uint32_t));
    memset(h->entries, 0xFF, nsize * sizeof(uint32_t));
    h->size = nsize;
    h->mask = nsize - 1;
    for (size_t i = 0; i < osize; i++) {
      uint32_t k = old[i];
      if (k != HS_EMPTY) {
        int l = (int)(hs_hash(k) & h->mask);
        while (h->entries[l] != HS_EMPTY && h->entries[l] != k) l = (int)((l + 1) & h->mask);
        h->entries[l] = k;
        h->num_entries++;
      }
    }
    free(old);
  }
  if (h->entries[loc] != HS_EMPTY) { /* :39-44 */
    loc = (int)((loc + 1) & h->mask);
    while (h->entries[loc] != HS_EMPTY && h->entries[loc] != a) loc = (int)((loc + 1) & h->mask);
    if (h->entries[loc] == a) return 1;
  }
  h->entries[loc] = a;
  h->num_entries++;
  return 0;
}

oracle_hashset *oracle_hashset_new(long n) {
  oracle_hashset *h = (oracle_hashset *)malloc(sizeof(*h));
  hs_init(h, n);
  return h;
}
int oracle_hashset_test(oracle_hashset *h, uint32_t a) { return hs_test(h, a); }
long oracle_hashset_num_entries(const oracle_hashset *h) { return h->num_entries; }
long oracle_hashset_capacity(const oracle_hashset *h) { return (long)h->size; }
int oracle_hashset_grew(const oracle_hashset *h) { return h->grew; }
void oracle_hashset_free(oracle_hashset *h) {
  free(h->entries);
  free(h);
}

/* ------------------------------------------------------------------------ */
/* (id, dist) pairs and the reference's `less` (beamSearch.h:62-64)          */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t id;
  float dist;
} idd;

static inline int idd_less(idd a, idd b) {
  return a.dist < b.dist || (a.dist == b.dist && a.id < b.id);
}

static void idd_insertion_sort(idd *a, long n) {
  for (long i = 1; i < n; i++) {
    idd x = a[i];
    long j = i - 1;
    while (j >= 0 && idd_less(x, a[j])) {
      a[j + 1] = a[j];
      j--;
    }
    a[j + 1] = x;
  }
}

/* Any correct sort gives the reference's order: `less` is a strict total
 * order on distinct (id, dist) pairs. */
static void idd_sort(idd *a, long n) {
  while (n > 24) {
    /* median of three */
    long m = n / 2;
    idd x = a[0], y = a[m], z = a[n - 1], piv;
    if (idd_less(x, y)) piv = idd_less(y, z) ? y : (idd_less(x, z) ? z : x);
    else piv = idd_less(x, z) ? x : (idd_less(y, z) ? z : y);
    long i = 0, j = n - 1;
    for (;;) {
      while (idd_less(a[i], piv)) i++;
      while (idd_less(piv, a[j])) j--;
      if (i >= j) break;
      idd t = a[i];
      a[i] = a[j];
      a[j] = t;
      i++;
      j--;
    }
    /* recurse into the smaller side */
    long left = j + 1;
    if (left < n - left) {
      idd_sort(a, left);
      a += left;
      n -= left;
    } else {
      idd_sort(a + left, n - left);
      n = left;
    }
  }
  idd_insertion_sort(a, n);
}

/* std::unique with `a.first == b.first` (beamSearch.h:198-199, index.h:86-87) */
static long idd_unique_by_id(idd *a, long n) {
  if (n == 0) return 0;
  long w = 0;
  for (long r = 1; r < n; r++)
    if (a[r].id != a[w].id) a[++w] = a[r];
  return w + 1;
}

/* std::set_union(first1, last1, first2, last2, out, less) */
static long idd_set_union(const idd *a, long na, const idd *b, long nb, idd *out) {
  long i = 0, j = 0, o = 0;
  while (i < na) {
    if (j == nb) {
      while (i < na) out[o++] = a[i++];
      return o;
    }
    if (idd_less(b[j], a[i])) out[o++] = b[j++];
    else {
      out[o++] = a[i];
      if (!idd_less(a[i], b[j])) j++;
      i++;
    }
  }
  while (j < nb) out[o++] = b[j++];
  return o;
}

/* std::set_difference(first1, last1, first2, last2, out, less) */
static long idd_set_difference(const idd *a, long na, const idd *b, long nb, idd *out) {
  long i = 0, j = 0, o = 0;
  while (i < na && j < nb) {
    if (idd_less(a[i], b[j])) out[o++] = a[i++];
    else {
      if (!idd_less(b[j], a[i])) i++;
      j++;
    }
  }
  while (i < na) out[o++] = a[i++];
  return o;
}

/* ------------------------------------------------------------------------ */
/* filtered_beam_search (beamSearch.h:33-228), use_filtering == false        */
/* ------------------------------------------------------------------------ */

typedef struct {
  idd *frontier, *unvisited, *visited, *cand, *newf;
  uint32_t *pruned;
  long cap_f, cap_u, cap_n, cap_v, cap_c, cap_p;
} bs_scratch;

static void *grow(void *p, long *cap, long need, size_t elt) {
  if (need <= *cap) return p;
  long nc = *cap ? *cap : 64;
  while (nc < need) nc *= 2;
  *cap = nc;
  return realloc(p, (size_t)nc * elt);
}

static void bs_scratch_free(bs_scratch *S) {
  free(S->frontier);
  free(S->unvisited);
  free(S->visited);
  free(S->cand);
  free(S->newf);
  free(S->pruned);
  memset(S, 0, sizeof(*S));
}

typedef struct {
  long fsize, vsize;
  uint64_t dist_cmps, full_dist_cmps;
  int grew;
} bs_result;

/* The reference's software prefetches (graph.h:109-113, euclidian_point.h:135-139,
 * mips_point.h:94-98): every 64-byte line of a graph row or point row. */
static inline void prefetch_lines(const void *p, size_t bytes) {
  for (size_t o = 0; o < bytes; o += 64) __builtin_prefetch((const char *)p + o);
}

static int bs_core(bs_scratch *S, const oracle_graph *G, const oracle_points *P, const void *q,
                   long self_id, const uint32_t *starts, int nstarts, const oracle_qparams *QP,
                   bs_result *res) {
  long B = QP->beam_size;
  int max_degree = (int)QP->degree_limit; /* :49 */
  if (nstarts == 0 || nstarts > B) return -1; /* :51-57 */
  if (max_degree <= 0) return -1;
  struct oracle_hashset hs;
  hs_init(&hs, 2L * (10 + (int)B) * max_degree); /* :66 */

  long capf = 2 * (B > nstarts ? B : nstarts) + (long)G->max_degree + 8;
  S->frontier = (idd *)grow(S->frontier, &S->cap_f, capf, sizeof(idd));
  S->unvisited = (idd *)grow(S->unvisited, &S->cap_u, capf, sizeof(idd));
  S->newf = (idd *)grow(S->newf, &S->cap_n, capf, sizeof(idd));
  S->pruned = (uint32_t *)grow(S->pruned, &S->cap_p, (long)G->max_degree + 1, sizeof(uint32_t));

  idd *frontier = S->frontier;
  long fsize = 0;
  for (int s = 0; s < nstarts; s++) { /* :73-76 */
    frontier[fsize].id = starts[s];
    frontier[fsize].dist = pdist(P, starts[s], q);
    fsize++;
    hs_test(&hs, starts[s]);
  }
  idd_sort(frontier, fsize); /* :77 */
  memcpy(S->unvisited, frontier, (size_t)fsize * sizeof(idd)); /* :81-83 */
  long vsize = 0;
  uint64_t dist_cmps = (uint64_t)nstarts, full_dist_cmps = (uint64_t)nstarts; /* :90-91 */
  long remain = fsize; /* :92 */
  long num_visited = 0;
  long csize = 0;
  long offset = 0;

  while (remain > offset && num_visited < QP->limit) { /* :116 */
    idd current = S->unvisited[offset]; /* :119 */
    prefetch_lines(G->adj + (size_t)current.id * (G->max_degree + 1),
                   ((size_t)G->max_degree + 1) * 4); /* :124 */
    /* visited.insert at upper_bound (:125-126) */
    S->visited = (idd *)grow(S->visited, &S->cap_v, vsize + 1, sizeof(idd));
    long lo = 0, hi = vsize;
    while (lo < hi) {
      long mid = (lo + hi) / 2;
      if (idd_less(current, S->visited[mid])) hi = mid;
      else lo = mid + 1;
    }
    memmove(S->visited + lo + 1, S->visited + lo, (size_t)(vsize - lo) * sizeof(idd));
    S->visited[lo] = current;
    vsize++;
    num_visited++;
    int frontier_full = (fsize == B); /* :128 */

    /* :151-159 */
    const uint32_t *row = G->adj + (size_t)current.id * (G->max_degree + 1);
    long deg = (long)row[0];
    long num_elts = deg < QP->degree_limit ? deg : QP->degree_limit;
    long npruned = 0;
    for (long i = 0; i < num_elts; i++) {
      uint32_t a = row[1 + i];
      if (hs_test(&hs, a) || (self_id >= 0 && (long)a == self_id)) continue;
      prefetch_lines(row_of(P, a), (size_t)P->dim * (P->dtype == OR_F32 ? 4 : 1)); /* :157 */
      S->pruned[npruned++] = a;
    }
    dist_cmps += (uint64_t)npruned; /* :160 */

    float cutoff = frontier_full ? frontier[fsize - 1].dist : FLT_MAX; /* :173-175 */
    S->cand = (idd *)grow(S->cand, &S->cap_c, csize + npruned, sizeof(idd));
    for (long i = 0; i < npruned; i++) { /* :176-182 */
      uint32_t a = S->pruned[i];
      float dist = pdist(P, a, q);
      full_dist_cmps++;
      if (dist >= cutoff) continue;
      S->cand[csize].id = a;
      S->cand[csize].dist = dist;
      csize++;
    }
    /* :185-192 */
    if (csize == 0 || (QP->limit >= 2 * (int)B && (double)csize < QP->batch_factor * (int)B &&
                       offset + 1 < remain)) {
      offset++;
      continue;
    }
    offset = 0;
    idd_sort(S->cand, csize); /* :197 */
    long cu = idd_unique_by_id(S->cand, csize); /* :198-199 */
    long need = fsize + cu;
    S->frontier = (idd *)grow(S->frontier, &S->cap_f, need, sizeof(idd));
    S->unvisited = (idd *)grow(S->unvisited, &S->cap_u, need, sizeof(idd));
    S->newf = (idd *)grow(S->newf, &S->cap_n, need, sizeof(idd));
    frontier = S->frontier;
    long nsz = idd_set_union(frontier, fsize, S->cand, cu, S->newf); /* :202-205 */
    csize = 0; /* :206 */
    if (nsz > B) nsz = B; /* :209 */
    memcpy(frontier, S->newf, (size_t)nsz * sizeof(idd)); /* :212-214 */
    fsize = nsz;
    remain = idd_set_difference(frontier, fsize < B ? fsize : B, S->visited, vsize,
                                S->unvisited); /* :217-222 */
  }
  res->fsize = fsize;
  res->vsize = vsize;
  res->dist_cmps = dist_cmps;
  res->full_dist_cmps = full_dist_cmps;
  res->grew = hs.grew;
  free(hs.entries);
  return 0;
}

int oracle_beam_search(const oracle_graph *G, const oracle_points *P, const void *query,
                       long self_id, const uint32_t *starts, int nstarts, const oracle_qparams *QP,
                       uint32_t *frontier_ids, float *frontier_dists, int *frontier_size,
                       uint32_t *visited_ids, float *visited_dists, long *visited_size,
                       long visited_cap, uint64_t *dist_cmps, uint64_t *full_dist_cmps,
                       int *hash_grew) {
  bs_scratch S;
  memset(&S, 0, sizeof(S));
  bs_result r;
  int rc = bs_core(&S, G, P, query, self_id, starts, nstarts, QP, &r);
  if (rc == 0) {
    for (long i = 0; i < r.fsize; i++) {
      if (frontier_ids) frontier_ids[i] = S.frontier[i].id;
      if (frontier_dists) frontier_dists[i] = S.frontier[i].dist;
    }
    if (frontier_size) *frontier_size = (int)r.fsize;
    for (long i = 0; i < r.vsize && i < visited_cap; i++) {
      if (visited_ids) visited_ids[i] = S.visited[i].id;
      if (visited_dists) visited_dists[i] = S.visited[i].dist;
    }
    if (visited_size) *visited_size = r.vsize;
    if (dist_cmps) *dist_cmps = r.dist_cmps;
    if (full_dist_cmps) *full_dist_cmps = r.full_dist_cmps;
    if (hash_grew) *hash_grew = r.grew;
  }
  bs_scratch_free(&S);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* a tiny parallel_for over [0, n) with dynamic chunks (stands in for         */
/* parlay::parallel_for, beamSearch.h:636)                                   */
/* ------------------------------------------------------------------------ */

typedef void (*pf_body)(void *ctx, long i, int tid);
typedef struct {
  pf_body body;
  void *ctx;
  long n, chunk;
  atomic_long next;
  int tid;
} pf_job;
typedef struct {
  pf_job *job;
  int tid;
} pf_arg;

static void *pf_worker(void *p) {
  pf_arg *a = (pf_arg *)p;
  pf_job *j = a->job;
  for (;;) {
    long s = atomic_fetch_add(&j->next, j->chunk);
    if (s >= j->n) break;
    long e = s + j->chunk < j->n ? s + j->chunk : j->n;
    for (long i = s; i < e; i++) j->body(j->ctx, i, a->tid);
  }
  return NULL;
}

static void parallel_for(long n, long chunk, int nthreads, pf_body body, void *ctx) {
  if (nthreads < 1) nthreads = 1;
  if (n <= chunk || nthreads == 1) {
    for (long i = 0; i < n; i++) body(ctx, i, 0);
    return;
  }
  pf_job job;
  job.body = body;
  job.ctx = ctx;
  job.n = n;
  job.chunk = chunk;
  atomic_init(&job.next, 0);
  pthread_t th[256];
  pf_arg args[256];
  if (nthreads > 256) nthreads = 256;
  for (int t = 0; t < nthreads; t++) {
    args[t].job = &job;
    args[t].tid = t;
    if (t > 0) pthread_create(&th[t], NULL, pf_worker, &args[t]);
  }
  pf_worker(&args[0]);
  for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* qsearchAll + beam_search_rerank (beamSearch.h:528-646)                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  const oracle_graph *G;
  const oracle_points *P, *Q;
  const char *queries, *qqueries;
  size_t qstride, qqstride;
  uint32_t start;
  const oracle_qparams *QP;
  uint32_t *out_ids;
  float *out_dists;
  uint32_t *out_visited;
  uint64_t *out_cmps;
  int *out_grew;
  bs_scratch *scratch; /* per thread */
  atomic_int failed;
} sb_ctx;

static void sb_body(void *vctx, long i, int tid) {
  sb_ctx *c = (sb_ctx *)vctx;
  bs_scratch *S = &c->scratch[tid];
  const void *p = c->queries + (size_t)i * c->qstride;
  const oracle_points *SP = c->Q ? c->Q : c->P;
  const void *qp = c->Q ? (const void *)(c->qqueries + (size_t)i * c->qqstride) : p;
  bs_result r;
  long k = c->QP->k;
  if (bs_core(S, c->G, SP, qp, -1, &c->start, 1, c->QP, &r) != 0 || r.fsize < k) { /* :557-562 */
    atomic_store(&c->failed, 1);
    return;
  }
  if (c->out_visited) c->out_visited[i] = (uint32_t)r.vsize; /* :564-567 */
  if (c->out_cmps) c->out_cmps[i] = r.full_dist_cmps;
  if (c->out_grew) c->out_grew[i] = r.grew;
  uint32_t *oi = c->out_ids + (size_t)i * k;
  float *od = c->out_dists + (size_t)i * k;
  if (c->Q) { /* :569-587 */
    int num_check = (int)(c->QP->k * c->QP->rerank_factor);
    if (num_check > r.fsize) num_check = (int)r.fsize;
    idd *pts = S->newf; /* capacity >= fsize */
    for (int t = 0; t < num_check; t++) {
      uint32_t j = S->frontier[t].id;
      pts[t].id = j;
      pts[t].dist = pdist(c->P, j, p);
    }
    idd_sort(pts, num_check);
    for (long t = 0; t < k; t++) {
      oi[t] = pts[t].id;
      od[t] = pts[t].dist;
    }
  } else { /* :588-596 */
    for (long t = 0; t < k; t++) {
      uint32_t j = S->frontier[t].id;
      oi[t] = j;
      od[t] = pdist(c->P, j, p);
    }
  }
}

static size_t elt_size(int dtype) { return dtype == OR_F32 ? 4 : 1; }

int oracle_search_batch(const oracle_graph *G, const oracle_points *P, const oracle_points *Q,
                        const void *queries, const void *qqueries, size_t nq, uint32_t start,
                        const oracle_qparams *QP, int nthreads, uint32_t *out_ids,
                        float *out_dists, uint32_t *out_visited, uint64_t *out_cmps,
                        int *out_hash_grew) {
  if (QP->k > QP->beam_size) return -1; /* :612-616 */
  if (nthreads < 1) nthreads = 1;
  sb_ctx c;
  memset(&c, 0, sizeof(c));
  c.G = G;
  c.P = P;
  c.Q = Q;
  c.queries = (const char *)queries;
  c.qqueries = (const char *)qqueries;
  c.qstride = P->dim * elt_size(P->dtype);
  c.qqstride = Q ? Q->dim * elt_size(Q->dtype) : 0;
  c.start = start;
  c.QP = QP;
  c.out_ids = out_ids;
  c.out_dists = out_dists;
  c.out_visited = out_visited;
  c.out_cmps = out_cmps;
  c.out_grew = out_hash_grew;
  c.scratch = (bs_scratch *)calloc((size_t)nthreads, sizeof(bs_scratch));
  atomic_init(&c.failed, 0);
  parallel_for((long)nq, 8, nthreads, sb_body, &c);
  for (int t = 0; t < nthreads; t++) bs_scratch_free(&c.scratch[t]);
  free(c.scratch);
  return atomic_load(&c.failed) ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* robustPrune (index.h:63-137)                                              */
/* ------------------------------------------------------------------------ */

static int robust_prune_core(const uint32_t *adj, unsigned maxdeg, const oracle_points *P,
                             uint32_t p, idd *cands, long nc, double alpha, int R, int add,
                             uint32_t *out, long *dist_comps) {
  long comps = 0;
  const uint32_t *row = adj + (size_t)p * (maxdeg + 1);
  if (add && ORACLE_MUTANT != 5) { /* :72-77 */
    for (uint32_t i = 0; i < row[0]; i++) {
      comps++;
      cands[nc].id = row[1 + i];
      cands[nc].dist = oracle_distance(row_of(P, row[1 + i]), row_of(P, p), P->dim, P->dtype,
                                       P->metric);
      nc++;
    }
  }
  idd_sort(cands, nc); /* :83 */
  nc = idd_unique_by_id(cands, nc); /* :86-88 */
  int nout = 0;
  long idx = 0;
  while (nout < R && idx < nc) { /* :95-116 */
    uint32_t p_star = cands[idx].id;
    idx++;
    if (p_star == p || p_star == 0xFFFFFFFFu) continue;
    out[nout++] = p_star;
    for (long i = idx; i < nc; i++) {
      uint32_t p_prime = cands[i].id;
      if (p_prime != 0xFFFFFFFFu) {
        comps++;
        float d_sp = oracle_distance(row_of(P, p_star), row_of(P, p_prime), P->dim, P->dtype,
                                     P->metric);
        float d_pp = cands[i].dist;
        const int prune = ORACLE_MUTANT == 2 ? alpha * (double)d_sp < (double)d_pp
                                             : alpha * (double)d_sp <= (double)d_pp; /* :111 */
        if (prune) cands[i].id = 0xFFFFFFFFu;
      }
    }
  }
  if (dist_comps) *dist_comps = comps;
  return nout;
}

int oracle_robust_prune(const oracle_graph *G, const oracle_points *P, uint32_t p,
                        const uint32_t *cand_ids, const float *cand_dists, int ncand,
                        double alpha, int R, int add, uint32_t *out, long *dist_comps) {
  const uint32_t *row = G->adj + (size_t)p * (G->max_degree + 1);
  long cap = ncand + (add ? (long)row[0] : 0) + 1;
  idd *c = (idd *)malloc((size_t)cap * sizeof(idd));
  for (int i = 0; i < ncand; i++) {
    c[i].id = cand_ids[i];
    c[i].dist = cand_dists[i];
  }
  int n = robust_prune_core(G->adj, G->max_degree, P, p, c, ncand, alpha, R, add, out, dist_comps);
  free(c);
  return n;
}

/* ------------------------------------------------------------------------ */
/* build_index / batch_insert (index.h:150-319)                               */
/* ------------------------------------------------------------------------ */

typedef struct {
  const oracle_points *P;
  uint32_t *adj;
  unsigned R, L;
  double alpha, batch_factor;
  const uint32_t *order; /* shuffled inserts */
  size_t floor_;
  uint32_t *new_out;     /* batch x R */
  int *new_cnt;
  bs_scratch *scratch;
  idd **pbuf;            /* per-thread prune buffers */
  long *pbuf_cap;
  /* reverse-edge groups */
  uint32_t *gkeys;
  long *gstart;
  uint32_t *gvals;
  atomic_int failed;
} bi_ctx;

static idd *pbuf_get(bi_ctx *c, int tid, long need) {
  c->pbuf[tid] = (idd *)grow(c->pbuf[tid], &c->pbuf_cap[tid], need, sizeof(idd));
  return c->pbuf[tid];
}

static void bi_search_body(void *vctx, long bi, int tid) {
  bi_ctx *c = (bi_ctx *)vctx;
  size_t i = c->floor_ + (size_t)bi;
  uint32_t index = c->order[i];
  oracle_graph G = {c->adj, c->P->n, c->R};
  oracle_qparams QP = {0, (long)c->L, (long)c->P->n, (long)c->R, 100.0, c->batch_factor};
  uint32_t sp = 0; /* set_start (:148) */
  bs_result r;
  bs_scratch *S = &c->scratch[tid];
  if (bs_core(S, &G, c->P, row_of(c->P, index), (long)index, &sp, 1, &QP, &r) != 0) {
    atomic_store(&c->failed, 1);
    return;
  }
  /* robustPrune(index, visited, G, Points, alpha) (:262-267) */
  long cap = r.vsize + c->R + 1;
  idd *buf = pbuf_get(c, tid, cap);
  memcpy(buf, S->visited, (size_t)r.vsize * sizeof(idd));
  c->new_cnt[bi] = robust_prune_core(c->adj, c->R, c->P, index, buf, r.vsize, c->alpha,
                                     (int)c->R, 1, c->new_out + (size_t)bi * c->R, NULL);
}

static void bi_group_body(void *vctx, long g, int tid) {
  bi_ctx *c = (bi_ctx *)vctx;
  uint32_t index = c->gkeys[g];
  long s = c->gstart[g], e = c->gstart[g + 1];
  long nc = e - s;
  uint32_t *row = c->adj + (size_t)index * (c->R + 1);
  long newsize = nc + (long)row[0];
  if (ORACLE_MUTANT == 1 ? newsize < (long)c->R
                         : newsize <= (long)c->R) { /* :295-297 add_neighbors_without_repeats */
    uint32_t tmp[1024];
    long t = 0;
    if (ORACLE_MUTANT == 6) {
      for (uint32_t i = 0; i < row[0]; i++) tmp[t++] = row[1 + i];
      for (long i = 0; i < nc; i++) {
        int found = 0;
        for (uint32_t j = 0; j < row[0]; j++) found |= row[1 + j] == c->gvals[s + i];
        if (!found) tmp[t++] = c->gvals[s + i];
      }
    } else {
      for (long i = 0; i < nc; i++) tmp[t++] = c->gvals[s + i];
      for (uint32_t i = 0; i < row[0]; i++) {
        uint32_t x = row[1 + i];
        int found = 0;
        for (long j = 0; j < nc; j++)
          if (c->gvals[s + j] == x) {
            found = 1;
            break;
          }
        if (!found) tmp[t++] = x;
      }
    }
    row[0] = (uint32_t)t;
    for (long i = 0; i < t; i++) row[1 + i] = tmp[i];
  } else { /* :298-302 robustPrune(index, ids, ...) */
    idd *buf = pbuf_get(c, tid, nc + (long)row[0] + 1);
    for (long i = 0; i < nc; i++) {
      uint32_t cid = c->gvals[s + i];
      buf[i].id = cid;
      buf[i].dist = oracle_distance(row_of(c->P, cid), row_of(c->P, index), c->P->dim,
                                    c->P->dtype, c->P->metric);
    }
    uint32_t out[1024];
    int n = robust_prune_core(c->adj, c->R, c->P, index, buf, nc, c->alpha, (int)c->R, 1, out,
                              NULL);
    row[0] = (uint32_t)n;
    for (int i = 0; i < n; i++) row[1 + i] = out[i];
  }
}

typedef struct {
  uint32_t key, val;
  long pos;
} kv;
static int kv_cmp(const void *a, const void *b) {
  const kv *x = (const kv *)a, *y = (const kv *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos ? 1 : 0);
}

/* batch_insert's schedule (index.h:202-234, base 2, max_fraction 0.02):
 * batches [floors[i], ceilings[i]) of the m inserts into n points.  Returns
 * the number of batches; writes at most cap of them. */
long oracle_vamana_batches(size_t n, size_t m, size_t *floors, size_t *ceilings, long cap) {
  size_t count = 0;
  size_t inc = 0;
  long nb = 0;
  const double base = 2.0, max_fraction = 0.02;
  size_t max_batch_size = (size_t)(max_fraction * (double)(float)n); /* :206-207 */
  if (max_batch_size > 1000000ul) max_batch_size = 1000000ul;
  if (max_batch_size == 0) max_batch_size = n; /* :208-209 */
  while (count < m) { /* :223-234 */
    size_t floor_, ceiling;
    if (ORACLE_MUTANT == 3) {
      floor_ = 0;
      ceiling = m;
      count = m;
    } else if (ORACLE_MUTANT == 4) {
      floor_ = count;
      ceiling = count + 1;
      count += 1;
    } else if (pow(base, (double)inc) <= (double)max_batch_size) {
      floor_ = (size_t)pow(base, (double)inc) - 1;
      size_t c2 = (size_t)pow(base, (double)(inc + 1)) - 1;
      ceiling = c2 < m ? c2 : m;
      count = ceiling;
    } else {
      floor_ = count;
      ceiling = count + max_batch_size < m ? count + max_batch_size : m;
      count += max_batch_size;
    }
    if (nb < cap) {
      floors[nb] = floor_;
      ceilings[nb] = ceiling;
    }
    nb++;
    inc += 1;
  }
  return nb;
}

static int batch_insert(bi_ctx *c, size_t m, int nthreads) {
  const long nbat = oracle_vamana_batches(c->P->n, m, NULL, NULL, 0);
  size_t *floors = (size_t *)malloc((size_t)(nbat + 1) * sizeof(size_t));
  size_t *ceilings = (size_t *)malloc((size_t)(nbat + 1) * sizeof(size_t));
  oracle_vamana_batches(c->P->n, m, floors, ceilings, nbat);
  int rc = 0;
  for (long b = 0; b < nbat && rc == 0; b++) {
    const size_t floor_ = floors[b], ceiling = ceilings[b];
    long bsz = (long)(ceiling - floor_);
    c->floor_ = floor_;
    c->new_out = (uint32_t *)malloc((size_t)bsz * c->R * sizeof(uint32_t));
    c->new_cnt = (int *)calloc((size_t)bsz, sizeof(int));
    parallel_for(bsz, 4, nthreads, bi_search_body, c); /* :247-269 */
    if (atomic_load(&c->failed)) {
      free(c->new_out);
      free(c->new_cnt);
      rc = -1;
      break;
    }
    for (long bi = 0; bi < bsz; bi++) { /* :271-273 update_neighbors */
      uint32_t *row = c->adj + (size_t)c->order[floor_ + bi] * (c->R + 1);
      row[0] = (uint32_t)c->new_cnt[bi];
      memcpy(row + 1, c->new_out + (size_t)bi * c->R, (size_t)c->new_cnt[bi] * sizeof(uint32_t));
    }
    /* :281-285 flatten (ngh, index) pairs and group by key, values kept in
     * flattened order */
    long npairs = 0;
    for (long bi = 0; bi < bsz; bi++) npairs += c->new_cnt[bi];
    kv *pairs = (kv *)malloc((size_t)(npairs + 1) * sizeof(kv));
    long t = 0;
    for (long bi = 0; bi < bsz; bi++)
      for (int j = 0; j < c->new_cnt[bi]; j++) {
        pairs[t].key = c->new_out[(size_t)bi * c->R + j];
        pairs[t].val = c->order[floor_ + bi];
        pairs[t].pos = t;
        t++;
      }
    qsort(pairs, (size_t)npairs, sizeof(kv), kv_cmp);
    long ng = 0;
    for (long i = 0; i < npairs; i++)
      if (i == 0 || pairs[i].key != pairs[i - 1].key) ng++;
    c->gkeys = (uint32_t *)malloc((size_t)(ng + 1) * sizeof(uint32_t));
    c->gstart = (long *)malloc((size_t)(ng + 2) * sizeof(long));
    c->gvals = (uint32_t *)malloc((size_t)(npairs + 1) * sizeof(uint32_t));
    long g = 0;
    for (long i = 0; i < npairs; i++) {
      if (i == 0 || pairs[i].key != pairs[i - 1].key) {
        c->gkeys[g] = pairs[i].key;
        c->gstart[g] = i;
        g++;
      }
      c->gvals[i] = pairs[i].val;
    }
    c->gstart[ng] = npairs;
    free(pairs);
    parallel_for(ng, 16, nthreads, bi_group_body, c); /* :292-303 */
    free(c->gkeys);
    free(c->gstart);
    free(c->gvals);
    free(c->new_out);
    free(c->new_cnt);
  }
  free(floors);
  free(ceilings);
  return rc;
}

typedef struct {
  const oracle_points *P;
  uint32_t *adj;
  unsigned R;
} sort_ctx;

static void sort_body(void *vctx, long i, int tid) {
  (void)tid;
  sort_ctx *c = (sort_ctx *)vctx;
  uint32_t *row = c->adj + (size_t)i * (c->R + 1);
  idd tmp[1024];
  uint32_t deg = row[0];
  for (uint32_t j = 0; j < deg; j++) {
    tmp[j].id = row[1 + j];
    tmp[j].dist = oracle_distance(row_of(c->P, (size_t)i), row_of(c->P, row[1 + j]), c->P->dim,
                                  c->P->dtype, c->P->metric);
  }
  idd_sort(tmp, deg);
  for (uint32_t j = 0; j < deg; j++) row[1 + j] = tmp[j].id;
}

int oracle_build_vamana(const oracle_points *P, uint32_t *adj, unsigned R, unsigned L,
                        double alpha, int num_passes, double batch_factor, const uint32_t *perms,
                        int sort_neighbors, int nthreads) {
  if (R == 0 || R > 1000 || L == 0 || nthreads < 1) return -1;
  bi_ctx c;
  memset(&c, 0, sizeof(c));
  c.P = P;
  c.adj = adj;
  c.R = R;
  c.L = L;
  c.batch_factor = batch_factor;
  c.scratch = (bs_scratch *)calloc((size_t)nthreads, sizeof(bs_scratch));
  c.pbuf = (idd **)calloc((size_t)nthreads, sizeof(idd *));
  c.pbuf_cap = (long *)calloc((size_t)nthreads, sizeof(long));
  atomic_init(&c.failed, 0);
  int rc = 0;
  for (int pass = 0; pass < num_passes && rc == 0; pass++) { /* :173-178 */
    c.alpha = (pass == num_passes - 1) ? alpha : 1.0;
    c.order = perms + (size_t)pass * P->n;
    rc = batch_insert(&c, P->n, nthreads);
  }
  if (rc == 0 && sort_neighbors) { /* :180-185 */
    sort_ctx sc = {P, adj, R};
    parallel_for((long)P->n, 256, nthreads, sort_body, &sc);
  }
  for (int t = 0; t < nthreads; t++) {
    bs_scratch_free(&c.scratch[t]);
    free(c.pbuf[t]);
  }
  free(c.scratch);
  free(c.pbuf);
  free(c.pbuf_cap);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* FP64 brute-force groundtruth (compute_groundtruth.cpp:22-60, in double)   */
/* ------------------------------------------------------------------------ */

typedef struct {
  const oracle_points *P;
  const char *queries;
  size_t qstride;
  int k;
  uint32_t *ids;
  double *dists;
} gt_ctx;

static double dist_f64(const oracle_points *P, size_t j, const void *q) {
  const void *p = row_of(P, j);
  unsigned d = P->dim;
  double r = 0.0;
  for (unsigned i = 0; i < d; i++) {
    double a, b;
    if (P->dtype == OR_F32) {
      a = ((const float *)p)[i];
      b = ((const float *)q)[i];
    } else if (P->dtype == OR_U8) {
      a = ((const uint8_t *)p)[i];
      b = ((const uint8_t *)q)[i];
    } else {
      a = ((const int8_t *)p)[i];
      b = ((const int8_t *)q)[i];
    }
    if (P->metric == OR_L2) r += (a - b) * (a - b);
    else r += a * b;
  }
  return P->metric == OR_L2 ? r : -r;
}

static void gt_body(void *vctx, long qi, int tid) {
  (void)tid;
  gt_ctx *c = (gt_ctx *)vctx;
  const void *q = c->queries + (size_t)qi * c->qstride;
  int k = c->k;
  uint32_t *ids = c->ids + (size_t)qi * k;
  double *ds = c->dists + (size_t)qi * k;
  int cnt = 0;
  for (size_t j = 0; j < c->P->n; j++) {
    double d = dist_f64(c->P, j, q);
    if (cnt == k && !(d < ds[k - 1] || (d == ds[k - 1] && j < ids[k - 1]))) continue;
    int pos = cnt < k ? cnt : k - 1;
    while (pos > 0 && (d < ds[pos - 1] || (d == ds[pos - 1] && j < ids[pos - 1]))) {
      ds[pos] = ds[pos - 1];
      ids[pos] = ids[pos - 1];
      pos--;
    }
    ds[pos] = d;
    ids[pos] = (uint32_t)j;
    if (cnt < k) cnt++;
  }
}

int oracle_groundtruth_f64(const oracle_points *P, const void *queries, size_t nq, int k,
                           int nthreads, uint32_t *out_ids, double *out_dists) {
  if (k <= 0 || (size_t)k > P->n) return -1;
  gt_ctx c = {P, (const char *)queries, P->dim * elt_size(P->dtype), k, out_ids, out_dists};
  parallel_for((long)nq, 1, nthreads, gt_body, &c);
  return 0;
}

/* Exposed for the sort's own property test. */
void oracle_sort_pairs(uint32_t *ids, float *dists, long n) {
  idd *a = (idd *)malloc((size_t)(n + 1) * sizeof(idd));
  for (long i = 0; i < n; i++) {
    a[i].id = ids[i];
    a[i].dist = dists[i];
  }
  idd_sort(a, n);
  for (long i = 0; i < n; i++) {
    ids[i] = a[i].id;
    dists[i] = a[i].dist;
  }
  free(a);
}

/* ------------------------------------------------------------------------ */
/* HNSW (algorithms/HNSW/HNSW.hpp): search and batch construction, with the  */
/* intended QueryParams(ef, ef, limit, threshold_m) of search_layer (:1096;  */
/* the reference's literal 1.35 lands in `limit`).                           */
/* ------------------------------------------------------------------------ */

static oracle_graph hnsw_level(const oracle_hnsw *H, int l) {
  oracle_graph g = {H->adj[l], H->n, H->max_degree[l]};
  return g;
}

/* search_layer (:1088-1109): beam search on level l from eps; frontier in S */
static int hnsw_search_layer(bs_scratch *S, const oracle_hnsw *H, const oracle_points *P,
                             const void *q, long self_id, const uint32_t *eps, int neps,
                             long ef, int l, long limit, bs_result *r) {
  oracle_graph g = hnsw_level(H, l);
  oracle_qparams QP = {ef, ef, limit > 0 ? limit : (long)H->n, (long)H->max_degree[l], 100.0,
                       0.125};
  return bs_core(S, &g, P, q, self_id, eps, neps, &QP, r);
}

typedef struct {
  const oracle_hnsw *H;
  const oracle_points *P;
  const uint32_t *entrance;
  int nentrance;
  const char *queries;
  size_t qstride;
  int k, ef;
  long limit;
  uint32_t *out_ids;
  float *out_dists;
  int *out_count;
  uint64_t *out_cmps;
  uint32_t *out_visited;
  bs_scratch *scratch;
  atomic_int failed;
} hs_ctx;

static void hs_body(void *vctx, long i, int tid) {
  hs_ctx *c = (hs_ctx *)vctx;
  bs_scratch *S = &c->scratch[tid];
  const void *q = c->queries + (size_t)i * c->qstride;
  uint32_t ep[64];
  int nep = c->nentrance;
  memcpy(ep, c->entrance, (size_t)nep * sizeof(uint32_t));
  uint64_t cmps = 0, vis = 0;
  bs_result r;
  for (int l = c->H->top; l > 0; l--) { /* search_layer_to (:1545-1566) */
    if (hnsw_search_layer(S, c->H, c->P, q, -1, ep, nep, 1, l, 0, &r) != 0) {
      atomic_store(&c->failed, 1);
      return;
    }
    cmps += r.full_dist_cmps;
    vis += (uint64_t)r.vsize;
    ep[0] = S->frontier[0].id; /* eps.push_back(W[0].u) */
    nep = 1;
  }
  if (hnsw_search_layer(S, c->H, c->P, q, -1, ep, nep, c->ef, 0, c->limit, &r) != 0) {
    atomic_store(&c->failed, 1);
    return;
  }
  cmps += r.full_dist_cmps;
  vis += (uint64_t)r.vsize;
  long m = r.fsize;
  uint32_t *oi = c->out_ids + (size_t)i * c->ef;
  float *od = c->out_dists + (size_t)i * c->ef;
  for (long t = 0; t < m; t++) {
    oi[t] = S->frontier[t].id;
    od[t] = S->frontier[t].dist;
  }
  if (m > c->k) { /* sort by distance (already in (dist, id) order), cut with ties (:1590-1596) */
    long kk = c->k;
    if (c->k > 0) {
      float dk = od[c->k - 1];
      while (kk < m && !(dk < od[kk])) kk++;
    }
    m = kk;
  }
  c->out_count[i] = (int)m;
  if (c->out_cmps) c->out_cmps[i] = cmps;
  if (c->out_visited) c->out_visited[i] = (uint32_t)vis;
}

int oracle_hnsw_search(const oracle_hnsw *H, const oracle_points *P, const uint32_t *entrance,
                       int nentrance, const void *queries, size_t nq, int k, int ef, long limit,
                       int nthreads, uint32_t *out_ids, float *out_dists, int *out_count,
                       uint64_t *out_cmps, uint32_t *out_visited) {
  if (nentrance < 1 || nentrance > 64 || ef < 1 || k < 0) return -1;
  if (nthreads < 1) nthreads = 1;
  hs_ctx c;
  memset(&c, 0, sizeof(c));
  c.H = H;
  c.P = P;
  c.entrance = entrance;
  c.nentrance = nentrance;
  c.queries = (const char *)queries;
  c.qstride = (size_t)P->dim * elt_size(P->dtype);
  c.k = k;
  c.ef = ef;
  c.limit = limit;
  c.out_ids = out_ids;
  c.out_dists = out_dists;
  c.out_count = out_count;
  c.out_cmps = out_cmps;
  c.out_visited = out_visited;
  c.scratch = (bs_scratch *)calloc((size_t)nthreads, sizeof(bs_scratch));
  atomic_init(&c.failed, 0);
  parallel_for((long)nq, 16, nthreads, hs_body, &c);
  for (int t = 0; t < nthreads; t++) bs_scratch_free(&c.scratch[t]);
  free(c.scratch);
  return atomic_load(&c.failed) ? -1 : 0;
}

/* ---- construction: HNSW::HNSW (:748-807) and insert (:809-996) ---- */

typedef struct {
  const oracle_hnsw *H;
  const oracle_points *P;
  const uint32_t *levels;
  float alpha;
  unsigned efc;
  uint32_t bb;        /* first node of the batch */
  uint32_t entrance;
  int level_ep, lc;
  uint32_t *eps;      /* batch x efc */
  int *neps;
  uint32_t *nbh;      /* batch x max_degree[0] */
  int *nnbh;
  bs_scratch *scratch;
  atomic_int failed;
} hb_ctx;

/* descent to the node's level with ef = 1 (:895-904) */
static void hb_descend_body(void *vctx, long i, int tid) {
  hb_ctx *c = (hb_ctx *)vctx;
  uint32_t u = c->bb + (uint32_t)i;
  int lu = (int)c->levels[u];
  uint32_t *ep = c->eps + (size_t)i * c->efc;
  ep[0] = c->entrance; /* eps_u = entrance */
  int nep = 1;
  bs_result r;
  for (int l = c->level_ep; l > lu; l--) {
    if (hnsw_search_layer(&c->scratch[tid], c->H, c->P, row_of(c->P, u), (long)u, ep, nep, 1, l,
                          0, &r) != 0) {
      atomic_store(&c->failed, 1);
      return;
    }
    ep[0] = c->scratch[tid].frontier[0].id;
    nep = 1;
  }
  c->neps[i] = nep;
}

/* prune_heuristic (:430-484) on a frontier sorted by (dist, id) */
static int hnsw_prune(const oracle_points *P, const idd *cand, long nc, unsigned size, float alpha,
                      uint32_t *out) {
  int nsel = 0;
  for (long t = 0; t < nc && nsel < (int)size; t++) {
    const float d_cu = cand[t].dist * alpha;
    int pruned = 0;
    for (int s = 0; s < nsel; s++) {
      float d_cr = oracle_distance(row_of(P, out[s]), row_of(P, cand[t].id), P->dim, P->dtype,
                                   P->metric);
      if (ORACLE_MUTANT == 7 ? d_cr <= d_cu : d_cr < d_cu) { /* :473 */
        pruned = 1;
        break;
      }
    }
    if (!pruned) out[nsel++] = cand[t].id;
  }
  return nsel;
}

/* one level of insert: search with efc, select neighbours (:910-925) */
static void hb_level_body(void *vctx, long i, int tid) {
  hb_ctx *c = (hb_ctx *)vctx;
  uint32_t u = c->bb + (uint32_t)i;
  if (c->lc > (int)c->levels[u]) return;
  bs_scratch *S = &c->scratch[tid];
  uint32_t *ep = c->eps + (size_t)i * c->efc;
  bs_result r;
  if (hnsw_search_layer(S, c->H, c->P, row_of(c->P, u), (long)u, ep, c->neps[i], c->efc, c->lc, 0,
                        &r) != 0) {
    atomic_store(&c->failed, 1);
    return;
  }
  unsigned M = c->H->max_degree[c->lc]; /* get_threshold_m(l_c) */
  c->nnbh[i] = hnsw_prune(c->P, S->frontier, r.fsize, M, c->alpha,
                          c->nbh + (size_t)i * c->H->max_degree[0]);
  for (long t = 0; t < r.fsize; t++) ep[t] = S->frontier[t].id; /* eps_u = res (:940-944) */
  c->neps[i] = (int)r.fsize;
}

typedef struct {
  hb_ctx *b;
  uint32_t *gkeys, *gvals;
  long *gstart;
  idd **buf;
  long *cap;
} hr_ctx;

/* reverse edges (:952-983): append when they fit, else the m nearest */
static void hr_body(void *vctx, long g, int tid) {
  hr_ctx *c = (hr_ctx *)vctx;
  const oracle_hnsw *H = c->b->H;
  const oracle_points *P = c->b->P;
  int lc = c->b->lc;
  unsigned md = H->max_degree[lc];
  uint32_t pv = c->gkeys[g];
  long s = c->gstart[g], e = c->gstart[g + 1];
  uint32_t *row = H->adj[lc] + (size_t)pv * (md + 1);
  long deg = row[0], nadd = e - s;
  if (ORACLE_MUTANT == 8 ? deg + nadd >= (long)md : deg + nadd > (long)md) { /* :966 */
    c->buf[tid] = (idd *)grow(c->buf[tid], &c->cap[tid], deg + nadd, sizeof(idd));
    idd *cand = c->buf[tid];
    for (long k = 0; k < deg; k++) {
      cand[k].id = row[1 + k];
      cand[k].dist = oracle_distance(row_of(P, row[1 + k]), row_of(P, pv), P->dim, P->dtype,
                                     P->metric);
    }
    for (long k = 0; k < nadd; k++) {
      cand[deg + k].id = c->gvals[s + k];
      cand[deg + k].dist = oracle_distance(row_of(P, c->gvals[s + k]), row_of(P, pv), P->dim,
                                           P->dtype, P->metric);
    }
    idd_sort(cand, deg + nadd);
    row[0] = md;
    for (unsigned k = 0; k < md; k++)
      row[1 + k] = ORACLE_MUTANT == 9 ? cand[deg + nadd - 1 - k].id : cand[k].id;
  } else {
    for (long k = 0; k < nadd; k++) row[1 + deg + k] = c->gvals[s + k];
    row[0] = (uint32_t)(deg + nadd);
  }
}

/* HNSW::HNSW's insert batches (HNSW.hpp:776-792): [begins[i], ends[i]) after
 * the first node.  Returns the number of batches; writes at most cap. */
long oracle_hnsw_batches(uint32_t n, float batch_base, uint32_t *begins, uint32_t *ends,
                         long cap) {
  uint32_t size_limit = (uint32_t)((double)n * 0.02); /* n*0.02 (:778) */
  if (size_limit == 0) size_limit = 1; /* the reference never ends a batch of 0 */
  uint32_t bb = 0, be = 1;
  long nb = 0;
  while (be < n) { /* :779-782 */
    bb = be;
    uint32_t b1 = (uint32_t)ceilf((float)bb * batch_base) + (ORACLE_MUTANT == 10 ? 0 : 1);
    be = n;
    if (b1 < be) be = b1;
    if (bb + size_limit < be) be = bb + size_limit;
    if (nb < cap) {
      begins[nb] = bb;
      ends[nb] = be;
    }
    nb++;
  }
  return nb;
}

int oracle_build_hnsw(const oracle_hnsw *H, const oracle_points *P, const uint32_t *levels,
                      unsigned efc, float alpha, float batch_base, uint32_t *entrance_out,
                      int nthreads) {
  const uint32_t n = (uint32_t)P->n;
  if (n == 0) return 0;
  if (nthreads < 1) nthreads = 1;
  for (uint32_t i = 0; i < n; i++)
    if ((int)levels[i] > H->top) return -1;
  uint32_t entrance = 0; /* node 0 starts the index (:763-774) */
  uint32_t bb = 0, be = 1;
  hb_ctx c;
  memset(&c, 0, sizeof(c));
  c.H = H;
  c.P = P;
  c.levels = levels;
  c.alpha = alpha;
  c.efc = efc;
  c.scratch = (bs_scratch *)calloc((size_t)nthreads, sizeof(bs_scratch));
  atomic_init(&c.failed, 0);
  size_t cap_batch = 0;
  hr_ctx rc;
  memset(&rc, 0, sizeof(rc));
  rc.b = &c;
  rc.buf = (idd **)calloc((size_t)nthreads, sizeof(idd *));
  rc.cap = (long *)calloc((size_t)nthreads, sizeof(long));
  kv *edges = NULL;
  size_t cap_edges = 0;
  const long nbat = oracle_hnsw_batches(n, batch_base, NULL, NULL, 0);
  uint32_t *begins = (uint32_t *)malloc((size_t)(nbat + 1) * sizeof(uint32_t));
  uint32_t *ends = (uint32_t *)malloc((size_t)(nbat + 1) * sizeof(uint32_t));
  oracle_hnsw_batches(n, batch_base, begins, ends, nbat);
  for (long b = 0; b < nbat; b++) { /* :777-792 */
    bb = begins[b];
    be = ends[b];
    size_t sz = be - bb;
    if (sz > cap_batch) {
      cap_batch = sz;
      c.eps = (uint32_t *)realloc(c.eps, sz * efc * sizeof(uint32_t));
      c.neps = (int *)realloc(c.neps, sz * sizeof(int));
      c.nbh = (uint32_t *)realloc(c.nbh, sz * H->max_degree[0] * sizeof(uint32_t));
      c.nnbh = (int *)realloc(c.nnbh, sz * sizeof(int));
    }
    c.bb = bb;
    c.entrance = entrance;
    c.level_ep = (int)levels[entrance];
    parallel_for((long)sz, 8, nthreads, hb_descend_body, &c);
    for (int lc = c.level_ep; lc >= 0; lc--) { /* :907-985 */
      c.lc = lc;
      parallel_for((long)sz, 4, nthreads, hb_level_body, &c);
      if (atomic_load(&c.failed)) break;
      unsigned md = H->max_degree[lc];
      size_t ne = 0;
      for (size_t i = 0; i < sz; i++) { /* forward edges (:928-933) */
        uint32_t u = bb + (uint32_t)i;
        if (lc > (int)levels[u]) continue;
        uint32_t *row = H->adj[lc] + (size_t)u * (md + 1);
        row[0] = (uint32_t)c.nnbh[i];
        for (int t = 0; t < c.nnbh[i]; t++) row[1 + t] = c.nbh[i * H->max_degree[0] + t];
        ne += (size_t)c.nnbh[i];
      }
      if (ne > cap_edges) {
        cap_edges = ne;
        edges = (kv *)realloc(edges, ne * sizeof(kv));
        rc.gkeys = (uint32_t *)realloc(rc.gkeys, ne * sizeof(uint32_t));
        rc.gvals = (uint32_t *)realloc(rc.gvals, ne * sizeof(uint32_t));
        rc.gstart = (long *)realloc(rc.gstart, (ne + 1) * sizeof(long));
      }
      size_t t = 0;
      for (size_t i = 0; i < sz; i++) { /* edge_add (pv, pu) in batch order */
        uint32_t u = bb + (uint32_t)i;
        if (lc > (int)levels[u]) continue;
        for (int j = 0; j < c.nnbh[i]; j++) {
          edges[t].key = c.nbh[i * H->max_degree[0] + j];
          edges[t].val = u;
          edges[t].pos = (long)t;
          t++;
        }
      }
      qsort(edges, ne, sizeof(kv), kv_cmp); /* group_by_key, values in insertion order */
      long ng = 0;
      for (size_t j = 0; j < ne; j++) {
        if (j == 0 || edges[j].key != edges[j - 1].key) {
          rc.gkeys[ng] = edges[j].key;
          rc.gstart[ng] = (long)j;
          ng++;
        }
        rc.gvals[j] = edges[j].val;
      }
      if (rc.gstart) rc.gstart[ng] = (long)ne;
      parallel_for(ng, 16, nthreads, hr_body, &rc);
    }
    if (atomic_load(&c.failed)) break;
    uint32_t hi = bb; /* max_element: the first node of the highest level (:987-995) */
    for (uint32_t u = bb + 1; u < be; u++)
      if (levels[u] > levels[hi]) hi = u;
    if (levels[hi] > levels[entrance]) entrance = hi;
  }
  *entrance_out = entrance;
  for (int t = 0; t < nthreads; t++) {
    bs_scratch_free(&c.scratch[t]);
    free(rc.buf[t]);
  }
  free(c.scratch);
  free(rc.buf);
  free(rc.cap);
  free(c.eps);
  free(c.neps);
  free(c.nbh);
  free(c.nnbh);
  free(edges);
  free(rc.gkeys);
  free(rc.gvals);
  free(rc.gstart);
  free(begins);
  free(ends);
  return atomic_load(&c.failed) ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* The reference's 8-bit scalar quantization of float points                 */
/* (graph_index.cpp:86-99 builds these for its quant path)                   */
/* ------------------------------------------------------------------------ */

static int cmp_float(const void *a, const void *b) {
  float x = *(const float *)a, y = *(const float *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* parlay::kth_smallest_copy(s, k): the element of rank k (0-based) */
static float kth_smallest(const float *s, size_t n, size_t k) {
  float *c = (float *)malloc(n * sizeof(float));
  memcpy(c, s, n * sizeof(float));
  qsort(c, n, sizeof(float), cmp_float);
  float v = c[k < n ? k : n - 1];
  free(c);
  return v;
}

int oracle_q8_params_compute(const oracle_points *P, oracle_q8_params *out) {
  if (P->dtype != OR_F32 || P->n == 0) return -1;
  const size_t n = P->n;
  const unsigned d = P->dim;
  float *mins = (float *)malloc(n * sizeof(float));
  float *maxs = (float *)malloc(n * sizeof(float));
  int all_ints = 1;
  for (size_t i = 0; i < n; i++) {
    const float *x = (const float *)row_of(P, i);
    float mn = 0.0f, mx = 0.0f; /* sequences start at 0.0 (euclidian_point.h:224-226, mips_point.h:420-429) */
    for (unsigned j = 0; j < d; j++) {
      if (P->metric == OR_L2) /* ni[i] && x >= 0 && x - (long) x == 0 (:229) */
        all_ints = all_ints && x[j] >= 0 && (x[j] - (float)(long)x[j]) == 0;
      mn = x[j] < mn ? x[j] : mn;
      mx = x[j] > mx ? x[j] : mx;
    }
    mins[i] = mn;
    maxs[i] = mx;
  }
  memset(out, 0, sizeof(*out));
  if (P->metric == OR_L2) { /* Euclidian_Point<uint8_t>::generate_parameters (:219-243) */
    float min_val = mins[0], max_val = maxs[0];
    for (size_t i = 1; i < n; i++) {
      min_val = mins[i] < min_val ? mins[i] : min_val;
      max_val = maxs[i] > max_val ? maxs[i] : max_val;
    }
    if (all_ints) {
      if (max_val < 256) max_val = 255;
      min_val = 0;
    }
    out->slope = 255.0f / (max_val - min_val);            /* parameters(min, max, dims) (:104-107) */
    out->offset = (int32_t)roundf(min_val * out->slope);
  } else { /* Quantized_Mips_Point<8, true>::generate_parameters (:415-470), trimmed */
    const double cutoff = .0001;
    size_t min_rank = (size_t)(cutoff * (double)n);
    size_t max_rank = (size_t)((1.0 - cutoff) * (double)(n - 1));
    float min_val = kth_smallest(mins, n, min_rank);
    float max_val = kth_smallest(maxs, n, max_rank);
    float bound = max_val > -min_val ? max_val : -min_val;
    out->max_val = bound;
    out->scale = (float)(255 / 2) / bound;                  /* (range/2) / max_val (:304-305) */
  }
  free(mins);
  free(maxs);
  return 0;
}

void oracle_q8_translate(const oracle_q8_params *prm, int metric, const float *x, unsigned d,
                         void *codes) {
  if (metric == OR_L2) { /* Euclidian_Point<uint8_t>::translate_point (:188-214) */
    uint8_t *c = (uint8_t *)codes;
    if (prm->slope == 1.0f && prm->offset == 0) {
      for (unsigned j = 0; j < d; j++) c[j] = (uint8_t)(int)x[j];
    } else {
      for (unsigned j = 0; j < d; j++) {
        int64_t r = (int64_t)roundf(x[j] * prm->slope) - prm->offset;
        if (r < 0) r = 0;
        if (r > 255) r = 255;
        c[j] = (uint8_t)r;
      }
    }
  } else { /* Quantized_Mips_Point<8, true>::translate_point (:393-406) */
    int8_t *c = (int8_t *)codes;
    const float mv = prm->max_val, scale = prm->scale;
    for (unsigned j = 0; j < d; j++) {
      float pj = x[j];
      if (pj < -mv) c[j] = (int8_t)(-127);
      else if (pj > mv) c[j] = (int8_t)127;
      else c[j] = (int8_t)(int32_t)roundf(pj * scale);
    }
  }
}
