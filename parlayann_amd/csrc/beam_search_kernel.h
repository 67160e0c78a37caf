// beam_search_kernel.h -- wave-per-query beam search over a fixed-degree graph.
//
// Restates filtered_beam_search (beamSearch.h:33-228, use_filtering = false)
// for one query per 64-lane wave, replaying the reference's sequential rules
// exactly:
//   * one visit per step: the closest unvisited entry of the (possibly stale)
//     frontier (beamSearch.h:116-127);
//   * neighbours G[cur][0 .. min(deg, degree_limit)) not seen before and not
//     the query's own row are distance-compared (:151-160, :176-182).  With
//     exact counts the seen set is the reference's (an LDS hash with a per-wave
//     global overflow table that grows as hashset.h:12-55 does); by default
//     every neighbour is evaluated and the seen test against the frontier and
//     the pending candidates follows the distance, which leaves candidates,
//     merges and visits the reference's (the main loop says why);
//   * candidates with d >= cutoff (last frontier distance when the frontier is
//     full) are dropped (:173-181);
//   * the merge is deferred while |candidates| < batch_factor * beam and more
//     unvisited entries remain (:185-192), |candidates| counted before dedup;
//   * merge = sort by (dist, id), union with the frontier, trim to beam
//     (:197-222), done here as a rank merge in LDS.
// Results: the first k frontier entries (beam_search_rerank :588-596), or with
// a full-precision FP32 store the rerank branch (:569-587): FP32 distances of
// the first min(k * rerank_factor, |frontier|) entries, sorted, first k.
//
// Memory: one wave per workgroup, persistent grid pulling query indices from
// a device counter.  LDS per wave: two frontier buffers (u64 keys + visited
// flags), candidate buffers, the gather list and the visited hash.  The FP32
// query row for the rerank is staged in the candidate buffers once the search
// is over; it has a region of its own only for the in-order
// sum of rows past 8 lines, or when it is longer than those buffers.
#pragma once
#include "kernels.h"
#include "dist.h"

#include <float.h>

#include <atomic>
#include <mutex>

namespace pann {

struct WaveLds {
  uint64_t *F0, *F1;
  uint8_t *FV0, *FV1;
  uint64_t* C;
  uint64_t* Cs;
  uint32_t* NL;
  float* DL;
  float* QL;
  uint32_t* H;
  uint32_t* RC;
  uint32_t* FB;
  float* CT;  // FP8 residual store: q . c for every centroid
  // several waves per query: the command words of the cooperative launch
  // (command, three arguments, each helper wave's partial minimum)
  uint32_t* CO;
};

__host__ __device__ inline uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

__host__ __device__ inline uint32_t wave_lds_bytes(uint32_t fcap, uint32_t ccap, uint32_t dim,
                                                   uint32_t hslots_log2, uint32_t rc_log2,
                                                   uint32_t fb_log2, uint32_t ncent,
                                                   uint32_t q_own, uint32_t wpq) {
  uint32_t b = 0;
  b += 2 * align16(fcap * 8);
  b += 2 * align16(fcap);
  b += 2 * align16(ccap * 8);
  b += align16(128 * 4) * 2;
  b += q_own ? align16(dim * 4) : 0u;
  b += (1u << hslots_log2) * 4;
  b += rc_log2 ? (1u << rc_log2) * 4 : 0u;
  b += fb_log2 ? (1u << fb_log2) / 8 : 0u;
  b += align16(ncent * 4);
  b += wpq > 1 ? 64u : 0u;
  return b;
}

__device__ inline WaveLds carve(char* p, const SearchArgs& a) {
  WaveLds L;
  uint32_t o = 0;
  L.F0 = (uint64_t*)(p + o); o += align16(a.fcap * 8);
  L.F1 = (uint64_t*)(p + o); o += align16(a.fcap * 8);
  L.FV0 = (uint8_t*)(p + o); o += align16(a.fcap);
  L.FV1 = (uint8_t*)(p + o); o += align16(a.fcap);
  L.C = (uint64_t*)(p + o); o += align16(a.ccap * 8);
  L.Cs = (uint64_t*)(p + o); o += align16(a.ccap * 8);
  L.NL = (uint32_t*)(p + o); o += align16(128 * 4);
  L.DL = (float*)(p + o); o += align16(128 * 4);
  // without a region of its own the query row shares the candidate buffers C, Cs
  L.QL = a.q_own ? (float*)(p + o) : (float*)L.C;
  o += a.q_own ? align16(a.dim * 4) : 0u;
  L.H = (uint32_t*)(p + o);
  o += (1u << a.hslots_log2) * 4;
  L.RC = (uint32_t*)(p + o);
  o += a.rc_log2 ? (1u << a.rc_log2) * 4 : 0u;
  L.FB = (uint32_t*)(p + o);
  o += a.fb_log2 ? (1u << a.fb_log2) / 8 : 0u;
  L.CT = (float*)(p + o);
  o += align16(a.ncent * 4);
  L.CO = (uint32_t*)(p + o);
  return L;
}


// The LDS seen set: open addressing over buckets of 4 slots, probed bucket
// after bucket; one ds_read_b128 reads a whole bucket, so a lookup almost
// always resolves in one LDS access even for the slowest lane of the wave.
// Slots only ever go from empty to a key, so a key inserted into bucket b
// found every bucket before b full, and they stay full: a lookup that meets
// a bucket with an empty slot can stop.
__device__ __forceinline__ uint32_t bucket_of(uint32_t key, uint32_t log2_slots) {
  return (key * 0x9E3779B1u) >> (32u - (log2_slots - 2u));
}

__device__ __forceinline__ bool lds_contains(const uint32_t* H, uint32_t log2n, uint32_t key) {
  const uint32_t bmask = (1u << (log2n - 2u)) - 1u;
  uint32_t b = bucket_of(key, log2n);
  for (;;) {
    const uint4 v = *(const uint4*)(H + 4u * b);
    if (v.x == key || v.y == key || v.z == key || v.w == key) return true;
    if (v.x == PANN_SENTINEL || v.y == PANN_SENTINEL || v.z == PANN_SENTINEL ||
        v.w == PANN_SENTINEL)
      return false;
    b = (b + 1u) & bmask;
  }
}

// Exact insert-or-test.  Returns true when `key` was absent and is now in.
__device__ __forceinline__ bool lds_insert(uint32_t* H, uint32_t log2n, uint32_t key) {
  const uint32_t bmask = (1u << (log2n - 2u)) - 1u;
  uint32_t b = bucket_of(key, log2n);
  for (;;) {
    const uint4 v = *(const uint4*)(H + 4u * b);
    if (v.x == key || v.y == key || v.z == key || v.w == key) return false;
    int j = v.x == PANN_SENTINEL ? 0
            : v.y == PANN_SENTINEL ? 1
            : v.z == PANN_SENTINEL ? 2
            : v.w == PANN_SENTINEL ? 3 : -1;
    if (j < 0) {
      b = (b + 1u) & bmask;
      continue;
    }
    const uint32_t old = atomicCAS(&H[4u * b + (uint32_t)j], PANN_SENTINEL, key);
    if (old == PANN_SENTINEL) return true;
    if (old == key) return false;
    // another lane took this slot for a different key: read the bucket again
  }
}

// Open addressing with linear probing (the global overflow table).
__device__ __forceinline__ uint32_t hash_slot(uint32_t key, uint32_t log2n) {
  return (key * 0x9E3779B1u) >> (32u - log2n);
}

// Per-wave global table of 64-bit slots (tag << 32 | id).  A slot whose tag
// differs from this query's tag is empty, so the table is never cleared:
// every query of every launch carries a distinct tag.
__device__ __forceinline__ bool global_insert(unsigned long long* G, uint32_t log2n, uint32_t tag,
                                              uint32_t key) {
  const uint32_t mask = (1u << log2n) - 1u;
  const unsigned long long mine = ((unsigned long long)tag << 32) | key;
  uint32_t h = hash_slot(key, log2n);
  unsigned long long old = __hip_atomic_load(&G[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    if ((uint32_t)(old >> 32) != tag) {
      unsigned long long prev = atomicCAS(&G[h], old, mine);
      if (prev == old) return true;
      old = prev;  // another lane won this slot: look at it again
      continue;
    }
    if ((uint32_t)old == key) return false;
    h = (h + 1u) & mask;
    old = __hip_atomic_load(&G[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// index of the first entry of FV[start..f) that is 0 (unvisited), or -1
__device__ __forceinline__ int first_unvisited(const uint8_t* FV, int f, int start, int lane) {
  for (int t = start; t < f; t += PANN_WAVE) {
    int i = t + lane;
    uint64_t m = ballot(i < f && FV[i] == 0);
    if (m) return t + (int)__builtin_ctzll(m);
  }
  return -1;
}

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    int w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}

__device__ __forceinline__ bool global_contains(const unsigned long long* G, uint32_t log2n,
                                                uint32_t tag, uint32_t key) {
  const uint32_t mask = (1u << log2n) - 1u;
  uint32_t h = hash_slot(key, log2n);
  for (;;) {
    unsigned long long v = __hip_atomic_load(&G[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(v >> 32) != tag) return false;
    if ((uint32_t)v == key) return true;
    h = (h + 1u) & mask;
  }
}

// number of entries of sorted A[0..n) that are < x
__device__ __forceinline__ int lower_bound_u64(const uint64_t* A, int n, uint64_t x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (A[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t rc_slot(uint32_t id, uint32_t log2n) {
  return (id * 0x9E3779B1u) >> (32u - log2n);
}

// Frontier-id Bloom filter: two bits per id, from two multiplicative hashes
// other than the tables' own.
__device__ __forceinline__ uint32_t fb_h1(uint32_t id, uint32_t log2n) {
  return (id * 0x85EBCA77u) >> (32u - log2n);
}
__device__ __forceinline__ uint32_t fb_h2(uint32_t id, uint32_t log2n) {
  return ((id ^ 0x5BD1E995u) * 0xC2B2AE3Du) >> (32u - log2n);
}
__device__ __forceinline__ bool fb_test(const uint32_t* FB, uint32_t log2n, uint32_t id) {
  const uint32_t a = fb_h1(id, log2n), b = fb_h2(id, log2n);
  return ((FB[a >> 5] >> (a & 31u)) & (FB[b >> 5] >> (b & 31u)) & 1u) != 0u;
}

// Whether sorted A[0..n) holds xa, and xb (n >= 1, the same in every lane):
// two branch-free searches in step, so that their LDS reads overlap, each
// ending on the last entry <= its key (or on A[0])
__device__ __forceinline__ void sorted_contains2(const uint64_t* A, int n, uint64_t xa,
                                                 uint64_t xb, bool& ha, bool& hb) {
  int ba = 0, bb = 0;
  for (int len = n; len > 1;) {
    const int half = len >> 1;
    const uint64_t ya = A[ba + half], yb = A[bb + half];
    ba = ya <= xa ? ba + half : ba;
    bb = yb <= xb ? bb + half : bb;
    len -= half;
  }
  ha = A[ba] == xa;
  hb = A[bb] == xb;
}

// Distances of rows NL[0..ncmp) into DL[0..ncmp) with the search's store.
// PAD: NL is this wave's whole 128-entry list (not a cooperative share), so
// the gather may fill its tail (eval_group_t).
template <int ST, int MT, int NCH, bool SEQ, bool PAD = false>
__device__ __forceinline__ void eval_rows_at(const GroupDist<ST, MT, NCH>& D, const SearchArgs& a,
                                             const float* QL, uint32_t* NL, float* DL,
                                             int ncmp, int lane) {
  if constexpr (SEQ) {
    // rows of up to 8 lines: the group gather with the in-order lane chain and the
    // query in registers; longer rows (NCH 8 only): 8 lines at a time, query in LDS
    if (NCH < 8 || a.row_bytes <= 8u * 128u)
      eval_group_inorder<MT, NCH>(D, a.pts, a.pts_stride, (int)a.row_bytes, NL, ncmp, DL, lane);
    else
      eval_group_inorder_lds<MT, 8, 1>(QL, (int)a.dim, a.pts, a.pts_stride, NL, ncmp, DL, lane);
  } else {
    eval_group<ST, MT, NCH, PAD>(D, a.pts, a.pts_stride, (int)a.row_bytes,
                                 a.lane_words == 3u ? 3 : 4, NL, ncmp, DL, lane);
  }
}

template <int ST, int MT, int NCH, bool SEQ>
__device__ __forceinline__ void eval_rows(const GroupDist<ST, MT, NCH>& D, const SearchArgs& a,
                                          const WaveLds& L, int ncmp, int lane) {
  eval_rows_at<ST, MT, NCH, SEQ, true>(D, a, L.QL, L.NL, L.DL, ncmp, lane);
}

// FP32 distances of rows NL[0..m) of the full-precision store (the rerank),
// the reference's arithmetic.  The 8-lane gather and in-order lane chain in the
// kernels of the reduced-precision stores; one lane per row for the FP8 L2
// kernels (their query scales leave no registers for it) and for the FP32
// store's own kernels, whose rerank is rare and whose search slowed by 1.5 %
// with it compiled in.
template <int ST, int MT>
__device__ __forceinline__ void eval_rerank_rows(const SearchArgs& a, const float* QL,
                                                 const uint32_t* NL, float* DL, int m, int lane) {
  constexpr bool group_rr = ST != ST_F32 && !(ST == ST_FP8 && MT == M_L2);
  if (group_rr && a.dim <= 128u)
    eval_group_inorder_lds<MT, 4, 1>(QL, (int)a.dim, a.full, a.full_stride, NL, m, DL, lane);
  else if (group_rr)
    eval_group_inorder_lds<MT, 8, 1>(QL, (int)a.dim, a.full, a.full_stride, NL, m, DL, lane);
  else
    eval_seq_f32<MT>(QL, a.full, a.full_stride, (int)a.dim, NL, m, DL, lane);
}

// ---------------------------------------------------------------------------
// The merge (beamSearch.h:197-222) as two phases over threads tid of nt, so
// that the waves of a cooperative launch share it: rank-sort the candidates
// C[0..c) into Cs, then place Cs and the frontier by rank into the other
// frontier buffer (first min(B, f + c) positions).  Keys are unique, so every
// rank is a distinct position.  merge_place returns the lowest new position
// of an unvisited entry among this thread's entries (INT_MAX if none).
// ---------------------------------------------------------------------------
// The rank of C[j]: the candidates smaller than it, read 8 at a time, so that
// a rank costs c / 8 LDS round trips rather than c.
__device__ __forceinline__ void merge_rank_sort(const WaveLds& L, int c, int tid, int nt) {
  for (int j = tid; j < c; j += nt) {
    const uint64_t x = L.C[j];
    int r = 0;
    int i = 0;
    for (; i + 8 <= c; i += 8) {
      uint64_t y[8];
#pragma unroll
      for (int t = 0; t < 8; t++) y[t] = L.C[i + t];
#pragma unroll
      for (int t = 0; t < 8; t++) r += (y[t] < x);
    }
    for (; i < c; i++) r += (L.C[i] < x);
    L.Cs[r] = x;
  }
}

__device__ __forceinline__ int merge_place(const WaveLds& L, int cur_buf, int c, int f, int B,
                                           int tid, int nt) {
  const uint64_t* F = cur_buf ? L.F1 : L.F0;
  const uint8_t* FV = cur_buf ? L.FV1 : L.FV0;
  uint64_t* F2 = cur_buf ? L.F0 : L.F1;
  uint8_t* FV2 = cur_buf ? L.FV0 : L.FV1;
  int first = 0x7FFFFFFF;
  for (int j = tid; j < c; j += nt) {
    const uint64_t x = L.Cs[j];
    const int pos = j + lower_bound_u64(F, f, x);
    if (pos < B) {
      F2[pos] = x;
      FV2[pos] = 0;
      first = pos < first ? pos : first;
    }
  }
  // the frontier: each thread a contiguous run of entries, whose first one a
  // search places among Cs; the rest follow it by merging, Cs[p] the first
  // candidate not below the current entry (one LDS round trip per candidate
  // passed, rather than a search per entry)
  const int per = (f + nt - 1) / nt;
  const int i0 = tid * per;
  const int i1 = i0 + per < f ? i0 + per : f;
  if (i0 < i1) {
    int p = lower_bound_u64(L.Cs, c, F[i0]);
    uint64_t y = p < c ? L.Cs[p] : ~0ull;
    for (int i = i0; i < i1; i += 4) {
      uint64_t x[4];
      uint8_t v[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        x[t] = i + t < i1 ? F[i + t] : 0ull;
        v[t] = i + t < i1 ? FV[i + t] : (uint8_t)1;
      }
#pragma unroll
      for (int t = 0; t < 4; t++) {
        if (i + t < i1) {
          while (y < x[t]) {
            p++;
            y = p < c ? L.Cs[p] : ~0ull;
          }
          const int pos = i + t + p;
          if (pos < B) {
            F2[pos] = x[t];
            FV2[pos] = v[t];
            if (!v[t]) first = pos < first ? pos : first;
          }
        }
      }
    }
  }
  return first;
}

// The final rank of the re-scored rows R[0..nchk) (beamSearch.h:579-584): the
// first k go to oi / od.
__device__ __forceinline__ void rerank_rank(const uint64_t* R, int nchk, int k, uint32_t* oi,
                                            float* od, int tid, int nt) {
  for (int j = tid; j < nchk; j += nt) {
    const uint64_t x = R[j];
    int r = 0;
    for (int i = 0; i < nchk; i++) r += (R[i] < x);
    if (r < k) {
      oi[r] = key_id(x);
      od[r] = key_dist(x);
    }
  }
}

// ---------------------------------------------------------------------------
// Several waves per query (SearchArgs::wpq = W > 1): a workgroup of W waves
// takes one query.  Wave 0 runs the search as a single wave would; at each
// row gather, merge and rerank it posts a command in LDS (CO[0], arguments
// CO[1..3]) and every wave of the workgroup does a share of it between two
// workgroup barriers.  The other waves loop on commands until CO_EXIT, so
// every wave passes the same number of barriers.
// ---------------------------------------------------------------------------
enum : uint32_t {
  CO_QUERY = 1,        // load query row CO[1] into the gather registers
  CO_EVAL = 2,         // distances of NL[0..CO[1]) with the search's store
  CO_EVAL_RR = 3,      // FP32 distances of NL[0..CO[1]) (the rerank)
  CO_MERGE_SORT = 4,   // merge phase 1 over CO[1] candidates
  CO_MERGE_PLACE = 5,  // merge phase 2: c = CO[1], f = CO[2], cur_buf | B << 1 = CO[3]
  CO_RR_RANK = 6,      // rerank rank: nchk = CO[1], cur_buf | k << 1 = CO[2], query CO[3]
  CO_EXIT = 7
};

// rows of a gather taken by each wave: a multiple of the 8 rows one gather
// instruction covers
__device__ __forceinline__ int coop_share(int n, int W) { return ((n + W - 1) / W + 7) & ~7; }

// wave 0 posts a command; the first barrier of its two
__device__ __forceinline__ void coop_post(const WaveLds& L, int lane, uint32_t cmd, uint32_t a0,
                                          uint32_t a1, uint32_t a2) {
  if (lane == 0) {
    L.CO[0] = cmd;
    L.CO[1] = a0;
    L.CO[2] = a1;
    L.CO[3] = a2;
  }
  __syncthreads();
}

// Distances of NL[0..n) into DL[0..n): by this wave alone (W = 1), or shared
// by the workgroup's W waves, this wave taking the first share; rr: the FP32
// rerank.  The cooperative launch is an instantiation of its own (CO), so that
// the one-wave kernels carry none of it.
template <int ST, int MT, int NCH, bool SEQ, bool CO>
__device__ __forceinline__ void eval_n(const GroupDist<ST, MT, NCH>& D, const SearchArgs& a,
                                       const WaveLds& L, int W, int n, bool rr, int lane) {
  if (CO && W > 1) {
    coop_post(L, lane, rr ? CO_EVAL_RR : CO_EVAL, (uint32_t)n, 0u, 0u);
    const int per = coop_share(n, W);
    const int cnt = n < per ? n : per;
    if (rr) eval_rerank_rows<ST, MT>(a, L.QL, L.NL, L.DL, cnt, lane);
    else eval_rows_at<ST, MT, NCH, SEQ>(D, a, L.QL, L.NL, L.DL, cnt, lane);
    __syncthreads();
  } else if (rr) {
    eval_rerank_rows<ST, MT>(a, L.QL, L.NL, L.DL, n, lane);
  } else {
    eval_rows<ST, MT, NCH, SEQ>(D, a, L, n, lane);
  }
}

// One beam search of a query: the graph it walks, its beam, k, limit and
// start points.  A launch runs the HNSW upper levels as passes at beam 1
// (not final: the nearest id is handed on, nothing is written) and then the
// main search (final: results and counters written).
struct Pass {
  const uint32_t* nbrs;
  uint32_t nbr_stride;
  int dl;              // neighbours read per row: min(degree_limit, max degree)
  int B, k;
  int64_t limit;
  const uint32_t* sp;  // start ids, or nullptr: the single id start1
  int ns;
  uint32_t start1;
  uint32_t pass;       // index of this pass within the query (seen-table tag)
  bool final;
  uint32_t vis0, cmps0;  // visits and evaluations of the earlier passes
};
struct PassOut {
  uint32_t best;  // id of the nearest frontier entry
  uint32_t vis, cmps;
  bool ok;
};

template <int ST, int MT, int NCH, bool SEQ, bool EX, bool FB, bool CO>
__device__ __forceinline__ PassOut search_one(const SearchArgs& a, const WaveLds& L,
                                              const GroupDist<ST, MT, NCH>& D, uint32_t q,
                                              int lane, unsigned long long* ghash, const Pass& P) {
  const uint32_t gtag = a.ghash_tag_base + q * (a.nup + 1u) + P.pass + 1u;
  const int B = P.B;
  const int k = P.k;
  const uint32_t self = a.self_ids ? a.self_ids[q] : PANN_SENTINEL;
  // |C| < batch_factor * B  <=>  |C| < ceil(batch_factor * B) for integer |C|
  const int thr = (int)ceil(a.batch_factor * (double)B);
  const bool defer_ok = P.limit >= 2 * (int64_t)B;
  const int dl = P.dl;
  const uint32_t hlog2 = a.hslots_log2;
  const uint32_t hslots = 1u << hlog2;
  const uint32_t* sp = P.sp;
  const int ns = P.ns;
  if (B < 1 || B > (int)a.fcap || k < 0 || k > B || (P.final && k > (int)a.out_stride) ||
      ns < 1 || ns > B) {  // per-query values outside what this launch was sized for
    if (lane == 0) atomicOr(a.status, PANN_ST_BAD_PARAMS);
    return PassOut{PANN_SENTINEL, P.vis0, P.cmps0, false};
  }

  // (the kernel loads the query once for all of its passes: into D's
  // registers, and into QL where the launch gives QL a region of its own)
  const uint64_t qr = a.self_ids ? (uint64_t)self : (uint64_t)q;  // build mode: the point's own row
  const char* qrow = a.queries + qr * a.q_stride;
  const float* fq = SEQ ? (const float*)qrow : (const float*)(a.fqueries + qr * a.fq_stride);
  // ---- clear the visited hash and the cache of evaluated ids ----
  for (uint32_t i = lane * 4; i < hslots; i += PANN_WAVE * 4)
    *(uint4*)(L.H + i) = make_uint4(PANN_SENTINEL, PANN_SENTINEL, PANN_SENTINEL, PANN_SENTINEL);
  const uint32_t rcl = EX ? 0u : a.rc_log2;
  // the filter is a separate instantiation (FB), so that kernels without it
  // keep their register count and occupancy
  const uint32_t fbl = FB ? a.fb_log2 : 0u;
  if (rcl)
    for (uint32_t i = lane * 4; i < (1u << rcl); i += PANN_WAVE * 4)
      *(uint4*)(L.RC + i) = make_uint4(PANN_SENTINEL, PANN_SENTINEL, PANN_SENTINEL, PANN_SENTINEL);
  wave_lds_sync();

  const int W = CO ? (int)a.wpq : 1;

  bool ovf = false;
  uint32_t lds_count = 0, g_count = 0;
  const uint32_t glog2 = a.ghash_log2;
  uint32_t status = 0;

  // ---- starting points (beamSearch.h:71-93) ----
  // evaluated 64 at a time into F1, then rank-sorted into F0 (equal keys keep
  // their order, so repeated starts stay repeated as in the reference)
  if (P.pass > 0) {
    // a later pass has one start, the last pass's nearest id, whose distance
    // (same store, same query) that pass left in DL[0]: not gathered again
    if (lane == 0) {
      if constexpr (EX) lds_insert(L.H, hlog2, P.start1);
      L.F1[0] = make_key(L.DL[0], P.start1);
    }
    wave_lds_sync();
  } else {
    for (int t = 0; t < ns; t += PANN_WAVE) {
      const int m = ns - t < PANN_WAVE ? ns - t : PANN_WAVE;
      if (lane < m) {
        const uint32_t s = sp ? sp[t + lane] : P.start1;
        L.NL[lane] = s;
        if constexpr (EX) lds_insert(L.H, hlog2, s);
      }
      wave_lds_sync();
      eval_n<ST, MT, NCH, SEQ, CO>(D, a, L, W, m, false, lane);
      wave_lds_sync();
      if (lane < m) L.F1[t + lane] = make_key(L.DL[lane], L.NL[lane]);
      wave_lds_sync();
    }
  }
  lds_count = (uint32_t)ns;
  int cur_buf = 0;
  for (int j = lane; j < ns; j += PANN_WAVE) {
    const uint64_t x = L.F1[j];
    int r = 0;
    for (int i = 0; i < ns; i++) {
      const uint64_t y = L.F1[i];
      r += (y < x) || (y == x && i < j);
    }
    L.F0[r] = x;
    L.FV0[r] = 0;
  }
  wave_lds_sync();
  int f = ns;
  uint32_t nvis = 0;
  uint32_t cmps = (uint32_t)ns;

  // The frontier-id filter: a Bloom filter of the ids in the frontier,
  // rebuilt whenever the frontier changes (here and after every merge).
  // A neighbour it reports is looked up exactly in F before its row is
  // gathered; one in F was seen before, so the reference's seen set skips it
  // and so does this search (its evaluation would only be discarded).
  auto fb_rebuild = [&](const uint64_t* Fx, int fx) {
    for (uint32_t i = lane; i < ((1u << fbl) >> 5); i += PANN_WAVE) L.FB[i] = 0u;
    wave_lds_sync();
    for (int i = lane; i < fx; i += PANN_WAVE) {
      const uint32_t id = key_id(Fx[i]);
      const uint32_t h1 = fb_h1(id, fbl), h2 = fb_h2(id, fbl);
      atomicOr(&L.FB[h1 >> 5], 1u << (h1 & 31u));
      atomicOr(&L.FB[h2 >> 5], 1u << (h2 & 31u));
    }
    wave_lds_sync();
  };
  // Which of two candidate lists (ids a where ca, ids b where cb) are in
  // F[0..f).  A wave-wide operation, called by all 64 lanes: only when the
  // filter reports something, each lane loads the ids of F[c + lane + 64 i]
  // (i < 8) for each 512-entry chunk c, and each filter hit still unresolved
  // costs one compare per register and a ballot.
  auto in_frontier2 = [&](const uint64_t* F, uint32_t ida, bool ca, uint32_t idb, bool cb,
                          bool& fa, bool& fbb) {
    uint64_t ma = ballot(ca && fb_test(L.FB, fbl, ida));
    uint64_t mb = ballot(cb && fb_test(L.FB, fbl, idb));
    fa = fbb = false;
    const uint32_t* Fw = (const uint32_t*)F;  // key_id = the low word of a key
    for (int c0 = 0; c0 < f && (ma | mb) != 0; c0 += 8 * PANN_WAVE) {
      uint32_t fid[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int t = c0 + lane + i * PANN_WAVE;
        fid[i] = t < f ? Fw[2 * t] : PANN_SENTINEL;
      }
      for (uint64_t m = ma; m; m &= m - 1) {
        const int j = __ffsll((long long)m) - 1;
        const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)ida, j);
        bool any = false;
#pragma unroll
        for (int i = 0; i < 8; i++) any |= fid[i] == x;
        if (ballot(any) != 0 && lane == j) fa = true;
      }
      for (uint64_t m = mb; m; m &= m - 1) {
        const int j = __ffsll((long long)m) - 1;
        const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)idb, j);
        bool any = false;
#pragma unroll
        for (int i = 0; i < 8; i++) any |= fid[i] == x;
        if (ballot(any) != 0 && lane == j) fbb = true;
      }
      ma &= ~ballot(fa);  // found: no later chunk needs to look
      mb &= ~ballot(fbb);
    }
  };
  if (fbl) fb_rebuild(L.F0, f);
  int c = 0;

  // ---- main loop (beamSearch.h:116-223) ----
  // u = the first unvisited frontier entry (unvisited_frontier[offset] of the
  // reference: entries before it are all visited).  The adjacency row of the
  // next candidate for the following step is loaded while this step gathers.
  int u = first_unvisited(L.FV0, f, 0, lane);
  uint32_t pref_id = PANN_SENTINEL;
  uint32_t nb_pref = PANN_SENTINEL;
  int acc_last = 0;  // candidates the last committed visit added to C
  // Frontier-and-pending seen set: an evaluated row with d < cutoff joins C
  // unless its key is already in the frontier F (binary search on the exact
  // (dist, id) key: a row's distance is computed the same way every time) or
  // already pending in C (the small LDS table H, emptied at every merge).  An
  // id the reference's seen set would also drop is dropped here.
  // test_rows: rows ra = lane of the first visit's n1 and rb = n1 + lane of the
  // second's nb2 (F and the cutoff stay the same until a merge, so both lists
  // are searched in F at once); ca / cb: under the cutoff and not in F.
  auto test_rows = [&](const uint64_t* F, int n1, int nb2, float cutoff, uint64_t& ka, bool& ca,
                       uint64_t& kb, bool& cb) {
    float da = FLT_MAX, db = FLT_MAX;
    ka = kb = 0;
    if (lane < n1) {
      da = L.DL[lane];
      ka = make_key(da, L.NL[lane]);
    }
    if (lane < nb2) {
      db = L.DL[n1 + lane];
      kb = make_key(db, L.NL[n1 + lane]);
    }
    const bool ua = lane < n1 && !(da >= cutoff), ub = lane < nb2 && !(db >= cutoff);
    ca = ua;
    cb = ub;
    // with the frontier-id filter on, ids in F were not gathered: none to find
    if (!fbl && ballot(ua || ub) != 0) {
      bool ha, hb;
      sorted_contains2(F, f, ka, kb, ha, hb);
      ca = ua && !ha;
      cb = ub && !hb;
    }
  };
  // commit_rows: a tested list's candidates join C (in lane order) unless
  // already pending; every row of the list (valid) enters the evaluated-id cache
  auto commit_rows = [&](uint64_t key, bool cand, bool valid) {
    const bool ins = cand && lds_insert(L.H, hlog2, key_id(key));
    if (rcl && valid) L.RC[rc_slot(key_id(key), rcl)] = key_id(key);  // evaluated and committed
    const uint64_t mi = ballot(ins);
    if (ins) L.C[c + (int)mbcnt(mi)] = key;
    c += __popcll(mi);
  };
  auto record_visit = [&](uint64_t k) {
    if (P.final && a.vis_out && lane == 0) {
      if (nvis < a.vis_cap) a.vis_out[(uint64_t)q * a.vis_cap + nvis] = k;
      else status |= PANN_ST_VISITED_OVERFLOW;
    }
    nvis++;
  };
  auto adjacency = [&](uint32_t id) -> uint32_t {
    return lane < dl ? P.nbrs[(uint64_t)id * P.nbr_stride + lane] : PANN_SENTINEL;
  };

  for (;;) {
    uint64_t* F = cur_buf ? L.F1 : L.F0;
    uint8_t* FV = cur_buf ? L.FV1 : L.FV0;
    if (u < 0 || (int64_t)nvis >= P.limit) break;
    const uint64_t cur_key = F[u];
    const uint32_t cur = key_id(cur_key);
    if (lane == 0) FV[u] = 1;
    record_visit(cur_key);
    const bool full = (f == B);
    const float cutoff = full ? key_dist(F[f - 1]) : FLT_MAX;

    // neighbours of cur (graph rows are padded with the sentinel past deg)
    const uint32_t nb = cur == pref_id ? nb_pref : adjacency(cur);
    // next unvisited entry of this (stale) frontier
    const int nu = first_unvisited(FV, f, u + 1, lane);

    if constexpr (EX) {
      // reference counting: every id enters the seen set at first encounter
      // (has_been_seen, :156), so dist_cmps equals the reference's
      pref_id = nu >= 0 ? key_id(F[nu]) : PANN_SENTINEL;
      nb_pref = nu >= 0 ? adjacency(pref_id) : PANN_SENTINEL;
      const bool valid = nb != PANN_SENTINEL;
      const int nvalid = __popcll(ballot(valid));
      // the LDS table takes ids up to half its slots, ids past it go to the
      // global table (at 7/8 load the longer probes made an HNSW 1M x 200
      // search 1.33x slower than at 1/2)
      if (!ovf && lds_count + (uint32_t)nvalid > (hslots >> 1)) ovf = true;
      // the global table also takes ids up to half its slots (the reference's
      // hashset grows past half load, hashset.h:25-38); an id it refuses
      // flags the launch, whose host call grows the table and searches again
      bool isnew = false, refused = false;
      if (valid) {
        if (!ovf) {
          isnew = lds_insert(L.H, hlog2, nb);
        } else if (!lds_contains(L.H, hlog2, nb)) {
          if (g_count + PANN_WAVE <= ((1u << glog2) >> 1)) {
            isnew = global_insert(ghash, glog2, gtag, nb);
          } else {
            isnew = true;
            refused = true;
          }
        }
      }
      const uint64_t mins = ballot(isnew);
      if (!ovf) lds_count += (uint32_t)__popcll(mins);
      else g_count += (uint32_t)__popcll(mins);
      if (ballot(refused) != 0) status |= PANN_ST_HASH_CAPACITY;
      const bool evalme = isnew && nb != self;  // Points[a].same_as(p) (:156)
      const uint64_t mnew = ballot(evalme);
      const int ncmp = __popcll(mnew);
      cmps += (uint32_t)ncmp;
      if (evalme) L.NL[mbcnt(mnew)] = nb;
      wave_lds_sync();
      eval_n<ST, MT, NCH, SEQ, CO>(D, a, L, W, ncmp, false, lane);
      wave_lds_sync();
      bool keep = false;
      uint64_t key = 0;
      if (lane < ncmp) {
        const float d = L.DL[lane];
        keep = !(d >= cutoff);
        key = make_key(d, L.NL[lane]);
      }
      const uint64_t mk = ballot(keep);
      if (keep) L.C[c + (int)mbcnt(mk)] = key;
      c += __popcll(mk);
      wave_lds_sync();
      // defer while the candidates are few and the stale frontier still has
      // unvisited entries (:185-192)
      if (c == 0 || (defer_ok && c < thr && nu >= 0)) {
        u = nu;
        continue;
      }
    } else {
      // Every neighbour is evaluated; the seen test happens after the
      // distance (test_rows, commit_rows).  Which ids the reference's seen set would
      // skip, and this kernel evaluates:
      //   * ids rejected earlier: the cutoff never rises once the frontier is
      //     full (and nothing is rejected before), so they are rejected again;
      //   * ids once accepted and since evicted from the frontier: eviction
      //     leaves the frontier full of B smaller keys, so they now fail the
      //     cutoff too;
      //   * ids in the frontier: skipped before their gather when the
      //     frontier-id filter is on, else found by test_rows;
      //   * ids pending in C: found by commit_rows.
      // Candidates, merges and visits are therefore the reference's; only the
      // count of distance evaluations exceeds its dist_cmps.
      //
      // Two visits per step: while merges are deferred the next visit is nu
      // with the same stale frontier and cutoff, so its neighbours are
      // gathered together with cur's.  Its results are committed only if the
      // first visit defers; its accepted ids are inserted after cur's, so an
      // id both lists share is taken once, for cur, as in the reference.
      // When cur's visit merges, the rows gathered for nu are discarded:
      // lookahead 2 gathers them only when the merge rule, applied to a
      // prediction of |C| after cur's visit, defers -- every neighbour
      // while the frontier is not full (there is no cutoff yet), else as
      // many as the last committed visit added.
      bool spec = a.lookahead == 1;
      if (a.lookahead == 2) {
        const int pred = c + (full ? acc_last : dl);
        spec = pred == 0 || (defer_ok && pred < thr);
      }
      const bool have2 = spec && nu >= 0 && (int64_t)nvis < P.limit;
      const uint64_t key2 = have2 ? F[nu] : 0;
      const uint32_t id2 = key_id(key2);
      const uint32_t nb2 = !have2 ? PANN_SENTINEL : (id2 == pref_id ? nb_pref : adjacency(id2));
      const int nnu = have2 ? first_unvisited(FV, f, nu + 1, lane) : -1;
      const int nxt = have2 ? nnu : nu;
      pref_id = nxt >= 0 ? key_id(F[nxt]) : PANN_SENTINEL;
      nb_pref = nxt >= 0 ? adjacency(pref_id) : PANN_SENTINEL;
      // an id found in the cache of committed evaluations was seen before:
      // the reference's seen set skips it, and its outcome here would be
      // the same (rejected again, or already in F or C), so it is not
      // evaluated again
      bool e1 = nb != PANN_SENTINEL && nb != self && !(rcl && L.RC[rc_slot(nb, rcl)] == nb);
      bool e2 = have2 && nb2 != PANN_SENTINEL && nb2 != self &&
                !(rcl && L.RC[rc_slot(nb2, rcl)] == nb2);
      if (fbl) {  // ids in the frontier were seen: not gathered again
        // a wave-wide call: every lane takes part, whatever its own flags
        bool f1, f2;
        in_frontier2(F, nb, e1, nb2, e2, f1, f2);
        e1 = e1 && !f1;
        e2 = e2 && !f2;
      }
      const uint64_t m1 = ballot(e1), m2 = ballot(e2);
      const int n1 = __popcll(m1), n2 = __popcll(m2);
      if (e1) L.NL[mbcnt(m1)] = nb;
      if (e2) L.NL[n1 + (int)mbcnt(m2)] = nb2;
      wave_lds_sync();
      eval_n<ST, MT, NCH, SEQ, CO>(D, a, L, W, n1 + n2, false, lane);
      wave_lds_sync();
      uint64_t ka, kb;
      bool ca, cb;
      test_rows(F, n1, n2, cutoff, ka, ca, kb, cb);
      const int c0 = c;
      commit_rows(ka, ca, lane < n1);
      acc_last = c - c0;
      cmps += (uint32_t)n1;
      wave_lds_sync();
      if (c == 0 || (defer_ok && c < thr && nu >= 0)) {
        if (!have2) {
          u = nu;
          continue;
        }
        if (lane == 0) FV[nu] = 1;
        record_visit(key2);
        const int c1 = c;
        commit_rows(kb, cb, lane < n2);
        acc_last = c - c1;
        cmps += (uint32_t)n2;
        wave_lds_sync();
        if (c == 0 || (defer_ok && c < thr && nnu >= 0)) {
          u = nnu;
          continue;
        }
      }
    }

    // ---- merge: rank-sort C, then place C and F by rank into the other buffer
    int first;  // lowest new position of an unvisited entry
    if (W > 1) {
      coop_post(L, lane, CO_MERGE_SORT, (uint32_t)c, 0u, 0u);
      merge_rank_sort(L, c, lane, PANN_WAVE * W);
      __syncthreads();
      coop_post(L, lane, CO_MERGE_PLACE, (uint32_t)c, (uint32_t)f,
                (uint32_t)cur_buf | ((uint32_t)B << 1));
      first = merge_place(L, cur_buf, c, f, B, lane, PANN_WAVE * W);
      __syncthreads();
      for (int w = 1; w < W; w++) {
        const int o = (int)L.CO[8 + w];
        first = o < first ? o : first;
      }
    } else {
      merge_rank_sort(L, c, lane, PANN_WAVE);
      wave_lds_sync();
      first = merge_place(L, cur_buf, c, f, B, lane, PANN_WAVE);
    }
    f = f + c < B ? f + c : B;
    c = 0;
    cur_buf ^= 1;
    if constexpr (!EX) {  // C is empty again: so is its table
      for (uint32_t i = lane * 4; i < hslots; i += PANN_WAVE * 4)
        *(uint4*)(L.H + i) = make_uint4(PANN_SENTINEL, PANN_SENTINEL, PANN_SENTINEL, PANN_SENTINEL);
    }
    u = wave_min_i32(first);
    if (u == 0x7FFFFFFF) u = -1;
    wave_lds_sync();
    if (fbl) fb_rebuild(cur_buf ? L.F1 : L.F0, f);
  }

  const uint64_t* F = cur_buf ? L.F1 : L.F0;
  if (!P.final) {  // an upper HNSW level: W[0], the nearest (HNSW.hpp:1557-1558)
    // and its distance in DL[0], where the next pass's start takes it
    const uint32_t best = key_id(F[0]);
    if (lane == 0) {
      L.DL[0] = key_dist(F[0]);
      if (status) atomicOr(a.status, status);
    }
    wave_lds_sync();
    return PassOut{best, P.vis0 + nvis, P.cmps0 + cmps, true};
  }
  // ---- results (beam_search_rerank :557-596) ----
  uint32_t* oi = a.out_ids + (uint64_t)q * a.out_stride;
  float* od = a.out_dists + (uint64_t)q * a.out_stride;
  if (f < k && !a.short_ok) status |= PANN_ST_SHORT_BEAM;
  if (a.full) {
    double nc = (double)k * a.rerank_factor;
    int nchk = nc < (double)f ? (int)nc : f;
    uint64_t* R = cur_buf ? L.F0 : L.F1;
    if (!a.q_own) {  // into the candidate buffers: no candidate is pending after the search
      for (uint32_t j = lane; j < a.dim; j += PANN_WAVE) L.QL[j] = fq[j];
      wave_lds_sync();
    }
    const uint64_t* S = F;  // the entries re-scored: F[0..nchk), or those kept in R
    if constexpr (ST == ST_FP8 || ST == ST_BF16 || (ST == ST_I8 && MT == M_IP)) {
      if (a.row_bound && nchk > k) {
        // Entries that cannot reach the first k keep their FP32 rows unread.
        // Each entry gets bounds lo <= g(D32) <= hi from its search distance d
        // and its row's bound, g the same increasing map for every row:
        //   FP8 or BF16, inner product: g = id, |d - D32| <= |q| b_x;
        //   the reference's 8-bit codes (inner product): d = -<code(q),
        //   code(x)> exactly, g(D) = s^2 D, |d - g(D32)| <= |s q| R_x +
        //   |s q - code(q)| N_x;
        //   FP8 or BF16, L2: g = sqrt; |sqrt(d) - |q - x^|| <= 2^-16 (sqrt(d) + |x^|)
        //   (the search sum's roundings), ||q - x^| - |q - x|| <= |x - x^| and
        //   sqrt(D32) is within eta = (dim + 8) 2^-24 relative of |q - x|,
        //   with b_x = |x - x^| + 2^-16 |x^|.
        // An entry whose lo lies above U = max over the first k of hi has
        // D32 above k smaller D32s.  A NaN among the first k's bounds makes U
        // infinite, and an entry with an infinite or NaN lo is kept: rows or
        // queries with non-finite values prune nothing.  The inner-product
        // sums carry a margin of 2^-22 of their magnitudes, which covers their
        // float roundings; the query norms are raised by 2^-16 for their own,
        // |s q - code(q)| also by 2^-20 |s q| for the rounding of s q.
        float qa = 0.0f, qb = 0.0f;
        if constexpr ((ST == ST_FP8 || ST == ST_BF16) && MT == M_IP) {
          for (uint32_t j = lane; j < a.dim; j += PANN_WAVE) qa += L.QL[j] * L.QL[j];
        } else if constexpr (ST == ST_I8) {
          const int8_t* cq = (const int8_t*)qrow;
          for (uint32_t j = lane; j < a.dim; j += PANN_WAVE) {
            const float v = a.q8_scale * L.QL[j], w = v - (float)cq[j];
            qa += v * v;
            qb += w * w;
          }
        }
        if constexpr (MT == M_IP) {
#pragma unroll
          for (int o = 32; o >= 1; o >>= 1) {
            qa += __shfl_xor(qa, o);
            if constexpr (ST == ST_I8) qb += __shfl_xor(qb, o);
          }
          qa = sqrtf(qa) * (1.0f + 0x1p-16f);
          qb = (sqrtf(qb) + 0x1p-20f * qa) * (1.0f + 0x1p-16f);
        }
        const float eta = (float)(a.dim + 8u) * 0x1p-24f;
        auto bounds = [&](uint64_t key, float& lo, float& hi) {
          const float d = key_dist(key);
          const uint32_t id = key_id(key);
          if constexpr (MT == M_L2) {
            // lo <= (sqrt(d) (1 - 2^-16) - b) (1 - eta) and hi >= (sqrt(d)
            // (1 + 2^-16) + b) (1 + eta): the factors carry 2^-16 and 2^-20
            // beyond those, more than the float roundings of sqrt, the
            // products and the one subtraction can take away
            const float u = d < 0.0f ? 0.0f : sqrtf(d), b = a.row_bound[id];
            lo = u * ((1.0f - 0x1p-15f) * (1.0f - 2.0f * eta)) - b * (1.0f + 0x1p-20f);
            hi = (u * (1.0f + 0x1p-15f) + b * (1.0f + 0x1p-20f)) * (1.0f + 2.0f * eta);
          } else {
            float e;
            if constexpr (ST == ST_FP8 || ST == ST_BF16) {
              e = qa * a.row_bound[id];
            } else {
              const float2 rn = *(const float2*)(a.row_bound + 2 * (uint64_t)id);
              e = qa * rn.x + qb * rn.y;
            }
            const float m = 0x1p-22f * (fabsf(d) + e);
            lo = d - e - m;
            hi = d + e + m;
          }
        };
        float U = -FLT_MAX;
        for (int i = lane; i < k; i += PANN_WAVE) {
          float lo, hi;
          bounds(F[i], lo, hi);
          U = fmaxf(U, hi == hi ? hi : HUGE_VALF);
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) U = fmaxf(U, __shfl_xor(U, o));
        int nk = 0;
        for (int t = 0; t < nchk; t += PANN_WAVE) {
          const int i = t + lane;
          bool keep = false;
          uint64_t key = 0;
          if (i < nchk) {
            key = F[i];
            float lo, hi;
            bounds(key, lo, hi);
            keep = i < k || !(lo > U && lo < HUGE_VALF);
          }
          const uint64_t mk = ballot(keep);
          if (keep) R[nk + (int)mbcnt(mk)] = key;
          nk += __popcll(mk);
        }
        wave_lds_sync();
        nchk = nk;
        S = R;
      }
    }
    for (int t = 0; t < nchk; t += PANN_WAVE) {
      int m = nchk - t < PANN_WAVE ? nchk - t : PANN_WAVE;
      if (lane < m) L.NL[lane] = key_id(S[t + lane]);
      wave_lds_sync();
      eval_n<ST, MT, NCH, SEQ, CO>(D, a, L, W, m, true, lane);
      wave_lds_sync();
      if (lane < m) R[t + lane] = make_key(L.DL[lane], L.NL[lane]);
      wave_lds_sync();
    }
    if (W > 1) {
      coop_post(L, lane, CO_RR_RANK, (uint32_t)nchk, (uint32_t)cur_buf | ((uint32_t)k << 1), q);
      rerank_rank(R, nchk, k, oi, od, lane, PANN_WAVE * W);
      __syncthreads();
    } else {
      rerank_rank(R, nchk, k, oi, od, lane, PANN_WAVE);
    }
  } else {
    for (int j = lane; j < k && j < f; j += PANN_WAVE) {
      oi[j] = key_id(F[j]);
      od[j] = key_dist(F[j]);
    }
  }
  if (lane == 0) {
    if (a.out_visited)
      a.out_visited[q] = (a.accumulate ? a.out_visited[q] : 0u) + P.vis0 + nvis;
    if (a.out_cmps) a.out_cmps[q] = (a.accumulate ? a.out_cmps[q] : 0u) + P.cmps0 + cmps;
    if (a.out_nfront) a.out_nfront[q] = (uint32_t)f;
    if (status) atomicOr(a.status, status);
  }
  wave_lds_sync();
  return PassOut{PANN_SENTINEL, P.vis0 + nvis, P.cmps0 + cmps, true};
}

// A helper wave of a cooperative launch: its share of each command wave 0
// posts, until CO_EXIT.
template <int ST, int MT, int NCH, bool SEQ>
__device__ __forceinline__ void coop_helper(const SearchArgs& a, const WaveLds& L, int lane, int wid,
                                            int W) {
  using QT = typename QueryElem<ST>::T;
  GroupDist<ST, MT, NCH> D;
  const int tid = wid * PANN_WAVE + lane, nt = PANN_WAVE * W;
  for (;;) {
    __syncthreads();
    const uint32_t cmd = L.CO[0], a0 = L.CO[1], a1 = L.CO[2], a2 = L.CO[3];
    if (cmd == CO_EXIT) break;
    if (cmd == CO_QUERY) {
      load_query_slices(D, (const QT*)(a.queries + (uint64_t)a0 * a.q_stride), a.scales,
                        (int)a.dim, lane & 7, a.lane_words == 3u ? 3 : 4);
      if constexpr (ST == ST_FP8R) D.T = L.CT;
    } else if (cmd == CO_EVAL || cmd == CO_EVAL_RR) {
      const int n = (int)a0, per = coop_share(n, W), lo = wid * per;
      const int cnt = n - lo < per ? n - lo : per;
      if (cnt > 0) {
        if (cmd == CO_EVAL)
          eval_rows_at<ST, MT, NCH, SEQ>(D, a, L.QL, L.NL + lo, L.DL + lo, cnt, lane);
        else
          eval_rerank_rows<ST, MT>(a, L.QL, L.NL + lo, L.DL + lo, cnt, lane);
      }
    } else if (cmd == CO_MERGE_SORT) {
      merge_rank_sort(L, (int)a0, tid, nt);
    } else if (cmd == CO_MERGE_PLACE) {
      int first = merge_place(L, (int)(a2 & 1u), (int)a0, (int)a1, (int)(a2 >> 1), tid, nt);
      first = wave_min_i32(first);
      if (lane == 0) L.CO[8 + wid] = (uint32_t)first;
    } else if (cmd == CO_RR_RANK) {
      const uint64_t* R = (a1 & 1u) ? L.F0 : L.F1;
      rerank_rank(R, (int)a0, (int)(a1 >> 1), a.out_ids + (uint64_t)a2 * a.out_stride,
                  a.out_dists + (uint64_t)a2 * a.out_stride, tid, nt);
    }
    __syncthreads();
  }
}

// Kernels that land just past 128 VGPRs are held to 128 (4 waves per SIMD, what their
// LDS allows at the beams they serve): FP8 inner-product rows of one or two lines (the
// 8-lane rerank takes them to 129-132), BF16 inner-product rows of two lines (96-128
// dimensions), one-line 8-bit rows and three-line FP32 rows, whose counts move
// around 128 with small changes elsewhere in the search.
// CO: workgroups of 2 or 4 waves per query (SearchArgs::wpq); otherwise one.
template <int ST, int MT, int NCH, bool SEQ, bool EX, bool FB, bool CO>
__global__ __launch_bounds__(CO ? 4 * PANN_WAVE : PANN_WAVE,
                             (!CO && !EX &&
                              ((MT == M_IP && ST == ST_FP8 && NCH == 1) ||
                               (MT == M_IP && ST == ST_FP8 && NCH == 2 && !FB) ||
                               (MT == M_IP && ST == ST_BF16 && NCH == 2) ||
                               ((ST == ST_I8 || ST == ST_U8) && NCH == 1 && !FB) ||
                               (ST == ST_F32 && !SEQ && NCH == 3))) ? 4 : 1)
void k_beam_search(SearchArgs a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  // (the one-wave kernels keep threadIdx.x itself: its known range [0, 64)
  // saves registers and scratch over the masked form)
  const int lane = CO ? ((int)threadIdx.x & (PANN_WAVE - 1)) : (int)threadIdx.x;
  const int W = CO ? (int)a.wpq : 1;
  WaveLds L = carve(smem, a);
  if constexpr (CO) {
    const int wid = (int)threadIdx.x / PANN_WAVE;
    if (wid > 0) {  // a helper wave of the cooperative launch
      coop_helper<ST, MT, NCH, SEQ>(a, L, lane, wid, W);
      return;
    }
  }
  unsigned long long* ghash = EX ? a.ghash + ((uint64_t)blockIdx.x << a.ghash_log2) : nullptr;
  for (;;) {
    uint32_t q = 0;
    if (lane == 0) q = atomicAdd(a.work_counter, 1u);
    q = readfirstlane(q);
    if (q >= a.nq) break;
    if (a.qorder) q = readfirstlane(a.qorder[q]);
    if (W > 1) {  // the helpers load the query row into their gather registers
      coop_post(L, lane, CO_QUERY, a.self_ids ? a.self_ids[q] : q, 0u, 0u);
      __syncthreads();
    }
    // the query, once for all of its passes: into registers (group mode), into
    // QL where the launch gives it a region of its own (sequential sums, the
    // rerank), and for the FP8 residual store its row of q . c into LDS
    GroupDist<ST, MT, NCH> D;
    {
      using QT = typename QueryElem<ST>::T;
      const uint64_t qr = a.self_ids ? (uint64_t)a.self_ids[q] : (uint64_t)q;
      const char* qrow = a.queries + qr * a.q_stride;
      load_query_slices(D, (const QT*)qrow, a.scales, (int)a.dim, lane & 7,
                        a.lane_words == 3u ? 3 : 4);
      if (a.q_own && (SEQ || a.full)) {
        const float* fq = SEQ ? (const float*)qrow : (const float*)(a.fqueries + qr * a.fq_stride);
        for (uint32_t j = lane; j < a.dim; j += PANN_WAVE) L.QL[j] = fq[j];
      }
      if constexpr (ST == ST_FP8R) {  // one per centroid, ncent a multiple of 4
        const float* row = a.qtable + qr * a.ncent;
        for (uint32_t c0 = 4u * lane; c0 < a.ncent; c0 += 4u * PANN_WAVE)
          *(float4*)(L.CT + c0) = *(const float4*)(row + c0);
        D.T = L.CT;
      }
      wave_lds_sync();
    }
    // passes 0 .. nup-1: the HNSW upper levels at beam 1 (search_layer_to,
    // HNSW.hpp:1545-1566), the first from the start points, each next one
    // from the last one's nearest id; pass nup: the search itself.  One call
    // site, so the search is inlined once.
    PassOut po{PANN_SENTINEL, 0u, 0u, true};
    for (uint32_t l = 0; l <= a.nup; l++) {
      const bool up = l < a.nup;
      Pass P;
      P.nbrs = up ? a.up[l].nbrs : a.nbrs;
      P.nbr_stride = up ? a.up[l].nbr_stride : a.nbr_stride;
      // an upper level: degree_limit = its max degree, limit = n
      P.dl = up ? (int)a.up[l].max_degree
                : (a.degree_limit < (int)a.max_degree ? a.degree_limit : (int)a.max_degree);
      P.B = readfirstlane(up ? 1 : (a.beam_pq ? a.beam_pq[q] : a.beam));
      P.k = readfirstlane(up ? 1 : (a.k_pq ? a.k_pq[q] : a.k));
      P.limit = up ? (int64_t)a.n : a.limit;
      if (l == 0) {
        P.sp = a.starts_pq ? a.starts_pq + (uint64_t)q * a.starts_stride : a.starts;
        P.ns = a.starts_pq ? (a.nstarts_pq ? (int)a.nstarts_pq[q] : 1) : (int)a.nstarts;
      } else {
        P.sp = nullptr;
        P.ns = 1;
      }
      P.start1 = po.best;
      P.pass = l;
      P.final = !up;
      P.vis0 = po.vis;
      P.cmps0 = po.cmps;
      po = search_one<ST, MT, NCH, SEQ, EX, FB, CO>(a, L, D, q, lane, ghash, P);
      po.best = readfirstlane(po.best);  // the same in every lane: keep it scalar
      po.vis = readfirstlane(po.vis);
      po.cmps = readfirstlane(po.cmps);
      if (!po.ok) break;
    }
  }
  if (W > 1) coop_post(L, lane, CO_EXIT, 0u, 0u, 0u);
}

template <int ST, int MT, int NCH, bool SEQ, bool EX, bool FB, bool CO>
static hipError_t launch_k(const SearchArgs& a, uint32_t grid, hipStream_t s) {
  uint32_t lds = search_lds_bytes(a);
  auto kern = k_beam_search<ST, MT, NCH, SEQ, EX, FB, CO>;
  // raise the kernel's dynamic-LDS limit only when a launch needs more, so a
  // repeated launch makes no attribute call (it may be under graph capture);
  // the check, the call and the record are one step under a lock, so two host
  // threads cannot leave the record above the attribute
  static std::atomic<uint32_t> lds_set{0};
  if (lds > lds_set.load()) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    if (lds > lds_set.load()) {
      hipError_t e = hipFuncSetAttribute((const void*)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
      lds_set.store(lds);
    }
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(PANN_WAVE * (CO ? a.wpq : 1u)), lds, s, a);
  return hipGetLastError();
}

// the exact-counts seen set and the frontier filter are instantiations of their
// own, so the default kernel carries none of their state; so is the
// cooperative launch (several waves per query), which takes neither
template <int ST, int MT, int NCH, bool SEQ>
static hipError_t launch_t(const SearchArgs& a, uint32_t grid, hipStream_t s) {
  if (a.wpq > 1) {
    if (a.exact_seen || a.fb_log2) return hipErrorInvalidValue;
    return launch_k<ST, MT, NCH, SEQ, false, false, true>(a, grid, s);
  }
  if (a.exact_seen) return launch_k<ST, MT, NCH, SEQ, true, false, false>(a, grid, s);
  return a.fb_log2 ? launch_k<ST, MT, NCH, SEQ, false, true, false>(a, grid, s)
                   : launch_k<ST, MT, NCH, SEQ, false, false, false>(a, grid, s);
}

#define PANN_NCH_CASES(ST, MT, SEQ)                              \
  switch (nch) {                                                 \
    case 1: return launch_t<ST, MT, 1, SEQ>(a, grid, s);         \
    case 2: return launch_t<ST, MT, 2, SEQ>(a, grid, s);         \
    case 3: return launch_t<ST, MT, 3, SEQ>(a, grid, s);         \
    case 4: return launch_t<ST, MT, 4, SEQ>(a, grid, s);         \
    case 8: return launch_t<ST, MT, 8, SEQ>(a, grid, s);         \
    default: return hipErrorInvalidValue;                        \
  }

// One store's launcher (both metrics, every row width); each store's kernels
// sit in a translation unit of their own, which the build compiles in parallel.
#define PANN_DEFINE_STORE_LAUNCH(NAME, ST, SEQ)                                          \
  hipError_t NAME(const SearchArgs& a, int metric, int nch, uint32_t grid, hipStream_t s) { \
    if (metric == M_L2) {                                                                \
      PANN_NCH_CASES(ST, M_L2, SEQ)                                                      \
    } else {                                                                             \
      PANN_NCH_CASES(ST, M_IP, SEQ)                                                      \
    }                                                                                    \
  }

}  // namespace pann
