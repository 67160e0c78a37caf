// pann_capi.cpp -- the C ABI of libpann.so (declared in include/pann.h).
//
// Host side of the device path: validates arguments where the reference
// aborts, lays the index out in HBM (rows padded to 128 bytes, graph rows
// padded with a sentinel past the degree), sizes the persistent launch and
// reports kernel status bits as pann_status codes.
#include "../../include/pann.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "dist.h"
#include "kernels.h"

namespace {

thread_local std::string g_err;

pann_status fail(pann_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

pann_status hip_fail(hipError_t e, const char* where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? PANN_ERR_OOM : PANN_ERR_DEVICE;
}

#define HIP_TRY(expr, where)                 \
  do {                                       \
    hipError_t _e = (expr);                  \
    if (_e != hipSuccess) return hip_fail(_e, where); \
  } while (0)

// Every device allocation the library makes goes through here, so that
// pann_device_alloc_count can show a caller that steady-state calls make none.
std::atomic<uint64_t> g_device_allocs{0};
template <typename T>
hipError_t dev_malloc(T** p, size_t bytes) {
  g_device_allocs.fetch_add(1, std::memory_order_relaxed);
  return hipMalloc((void**)p, bytes);
}

uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }
uint32_t ceil_log2(uint64_t x) {
  uint32_t l = 0;
  while ((1ull << l) < x) l++;
  return l;
}
size_t elem_bytes(int dt) { return dt == PANN_DT_F32 ? 4 : 1; }

// Sentinel-padded device copy of a host graph in Graph layout (graph.h:46-64).
pann_status upload_graph(uint32_t n, uint32_t md, uint32_t stride, const uint32_t* graph,
                         uint32_t* d_nbrs) {
  const uint64_t chunk = 1u << 20;
  std::vector<uint32_t> buf;
  for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
    uint64_t r1 = std::min<uint64_t>(n, r0 + chunk);
    buf.assign((r1 - r0) * stride, PANN_SENTINEL);
    if (graph)
      for (uint64_t i = r0; i < r1; i++) {
        const uint32_t* row = graph + i * (md + 1);
        std::memcpy(&buf[(i - r0) * stride], row + 1, row[0] * 4);
      }
    HIP_TRY(hipMemcpy(d_nbrs + r0 * stride, buf.data(), buf.size() * 4, hipMemcpyHostToDevice),
            "hipMemcpy(graph)");
  }
  return PANN_OK;
}

// QueryParams checks that hold for any batch, an empty one included: the
// reference's qsearchAll aborts on k > beam before looking at the queries.
pann_status check_qparams(const pann_qparams* qp) {
  const int64_t B = qp->beam_size;
  if (qp->k < 0 || B < 1) return fail(PANN_ERR_ARG, "beam_size must be >= 1 and k >= 0");
  if (qp->k > B)  // beamSearch.h:612-616
    return fail(PANN_ERR_ARG, "beam search parameter Q = " + std::to_string(B) +
                                  " same size or smaller than k = " + std::to_string(qp->k));
  if (B > 4096) return fail(PANN_ERR_UNSUPPORTED, "beam_size > 4096 is not supported");
  if (qp->degree_limit < 1) return fail(PANN_ERR_ARG, "degree_limit must be >= 1");
  if (!(qp->batch_factor >= 0.0) || !(qp->rerank_factor >= 0.0))
    return fail(PANN_ERR_ARG, "batch_factor and rerank_factor must be >= 0");
  return PANN_OK;
}

pann_status check_graph(uint32_t n, uint32_t md, const uint32_t* graph) {
  // graph.h:60-63 / :240-243 abort on bad degrees and ids; here they are errors
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t* row = graph + i * (md + 1);
    if (row[0] > md) return fail(PANN_ERR_ARG, "graph row " + std::to_string(i) + " exceeds max degree");
    for (uint32_t j = 0; j < row[0]; j++)
      if (row[1 + j] >= n)
        return fail(PANN_ERR_ARG, "graph row " + std::to_string(i) + " has an out-of-range id");
  }
  return PANN_OK;
}


}  // namespace

struct pann_index_s {
  int device = 0;
  uint32_t n = 0, dim = 0, max_degree = 0;
  int dtype = 0, metric = 0;
  char* d_pts = nullptr;
  uint64_t pts_stride = 0;
  uint32_t row_bytes = 0;
  char* d_bf16 = nullptr;
  uint64_t bf16_stride = 0;
  char* d_fp8 = nullptr;
  uint64_t fp8_stride = 0;
  float* d_scales = nullptr;
  uint32_t* d_absmax = nullptr;
  // per FP8 row, a bound on its FP8 distance's error (launch_fp8_row_bound), with
  // which the rerank skips rows that cannot reach k
  float* d_fp8_bound = nullptr;
  // per BF16 row, the same (launch_bf16_row_bound)
  float* d_bf16_bound = nullptr;
  // inner product: per Q8 row, (R, N) of launch_q8_row_bound, for the same purpose
  float* d_q8_bound = nullptr;
  float* d_xnorm = nullptr;        // |x|^2 per point (brute force, L2)
  // FP8 residual store (PANN_STORE_FP8_RESIDUAL): e4m3 codes of x - c with
  // per-dimension scales, c the nearest of ncent sampled points
  char* d_fp8r = nullptr;
  uint64_t fp8r_stride = 0;
  float* d_rscales = nullptr;
  uint32_t* d_rabsmax = nullptr;
  float* d_cent = nullptr;     // ncent x pts_stride (the sampled rows)
  uint32_t* d_assign = nullptr;  // n: each point's centroid
  uint32_t ncent = 0;
  char* d_q8 = nullptr;            // the reference's 8-bit codes (PANN_STORE_Q8)
  uint64_t q8_stride = 0;
  float q8_slope = 0, q8_max = 0, q8_scale = 0;
  int32_t q8_offset = 0;
  uint32_t* d_nbrs = nullptr;
  uint32_t nbr_stride = 0;
  // graph per level of a layered (HNSW) index; level 0 is d_nbrs above
  struct Level {
    uint32_t* d_nbrs = nullptr;
    uint32_t max_degree = 0, nbr_stride = 0;
  };
  std::vector<Level> levels;
  // Launch state per stream: launches on different streams may run at the
  // same time, so each stream has its own work counter, status word, start
  // points, scratch and events.  Created at a stream's first use (outside
  // any graph capture).
  struct StreamCtx {
    uint32_t* d_counter = nullptr;
    uint32_t* d_status = nullptr;
    uint32_t* d_starts = nullptr;
    std::vector<uint32_t> starts_host;  // what d_starts holds (copied only on change)
    unsigned long long* d_ghash = nullptr;
    size_t ghash_words = 0;
    uint32_t ghash_tag = 0;  // tags handed out so far (table zeroed when they wrap)
    uint64_t* d_vis = nullptr;
    size_t vis_words = 0;
    uint32_t* d_eps = nullptr;  // per-query entry points of the HNSW descent
    float* d_epsd = nullptr;
    size_t eps_cap = 0;
    uint8_t* d_qcodes = nullptr;  // per-search query codes for PANN_STORE_Q8
    size_t qcodes_cap = 0;
    float* d_qtable = nullptr;    // PANN_STORE_FP8_RESIDUAL: nq x ncent q . c
    size_t qtable_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the host-buffer entry points' device staging (queries in, results out),
    // grown geometrically so that repeated calls allocate nothing
    char* d_stage = nullptr;
    size_t stage_cap = 0;
    uint32_t* d_cmps = nullptr;  // build-mode search: evaluations per query
    size_t cmps_cap = 0;
    // per-query beams: the longest-first query order
    uint32_t* d_order = nullptr;
    size_t order_cap = 0;
  };
  std::unordered_map<hipStream_t, StreamCtx> streams;
  StreamCtx* last_timed = nullptr;
  uint32_t* d_dummy = nullptr;
  // distance evaluations of the build steps: [0] build-mode searches, [1]
  // robustPrune (pann_build_eval_counts)
  unsigned long long* d_evals = nullptr;
  uint64_t* d_pkeys = nullptr;   // prune sort scratch for groups too large for LDS
  uint8_t* d_pflags = nullptr;
  size_t pscratch_entries = 0;
  int num_cus = 0;
  int64_t lds_per_cu = 0;
  int64_t max_lds_block = 0;
  std::mutex mu;
  // serialises the host-buffer entry points, which share the default stream's
  // staging buffer
  std::mutex host_mu;
  // log2 growth of the exact seen set's global table: a synchronous search
  // that outgrows the table doubles it twice and runs again, and later
  // searches keep the larger table (ghash_log2_last: the last launch's size)
  std::atomic<uint32_t> ghash_grow{0};
  std::atomic<uint32_t> ghash_log2_last{0};
};

namespace {

// The overflow table holds (tag << 32 | id) slots; tag 0 never occurs, so a
// zeroed table is empty for every query.
using StreamCtx = pann_index_s::StreamCtx;

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive;
}

// The caller holds idx->mu.
pann_status stream_ctx(pann_index_t idx, hipStream_t s, StreamCtx** out) {
  auto it = idx->streams.find(s);
  if (it == idx->streams.end()) {
    if (capturing(s))
      return fail(PANN_ERR_ARG, "first use of a stream under graph capture: run the same "
                                "search once on that stream before capturing it");
    StreamCtx c;
    hipError_t e = dev_malloc(&c.d_counter, 256);
    if (e == hipSuccess) e = dev_malloc(&c.d_status, 256);
    if (e == hipSuccess) e = dev_malloc(&c.d_starts, 64 * 4);
    if (e == hipSuccess) e = hipEventCreate(&c.ev0);
    if (e == hipSuccess) e = hipEventCreate(&c.ev1);
    if (e != hipSuccess) {
      for (void* p : {(void*)c.d_counter, (void*)c.d_status, (void*)c.d_starts})
        if (p) (void)hipFree(p);
      if (c.ev0) (void)hipEventDestroy(c.ev0);
      if (c.ev1) (void)hipEventDestroy(c.ev1);
      return hip_fail(e, "stream context allocation");
    }
    it = idx->streams.emplace(s, std::move(c)).first;
  }
  *out = &it->second;
  return PANN_OK;
}

// The overflow table holds (tag << 32 | id) slots; tag 0 never occurs, so a
// zeroed table is empty for every query.
pann_status ensure_ghash(StreamCtx& C, size_t words, uint32_t nq) {
  if (C.ghash_words < words) {
    if (C.d_ghash) (void)hipFree(C.d_ghash);
    C.d_ghash = nullptr;
    C.ghash_words = 0;
    HIP_TRY(dev_malloc(&C.d_ghash, words * 8), "hipMalloc(visited overflow table)");
    HIP_TRY(hipMemset(C.d_ghash, 0, words * 8), "hipMemset(visited overflow table)");
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    C.ghash_words = words;
    C.ghash_tag = 0;
  }
  if ((uint64_t)C.ghash_tag + nq + 1 >= 0xFFFFFFFFull) {  // tags about to wrap
    HIP_TRY(hipMemset(C.d_ghash, 0, C.ghash_words * 8), "hipMemset(visited overflow table)");
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    C.ghash_tag = 0;
  }
  return PANN_OK;
}

// Staging of at least `bytes` for the host-buffer entry points: grown to
// twice the old size at least, never shrunk.  The caller holds idx->host_mu.
pann_status ensure_stage(StreamCtx& C, size_t bytes) {
  if (C.stage_cap >= bytes) return PANN_OK;
  const size_t cap = std::max<size_t>(bytes, 2 * C.stage_cap);
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");  // the old buffer may be in use
  if (C.d_stage) (void)hipFree(C.d_stage);
  C.d_stage = nullptr;
  C.stage_cap = 0;
  HIP_TRY(dev_malloc(&C.d_stage, cap), "hipMalloc(host staging)");
  C.stage_cap = cap;
  return PANN_OK;
}

pann_status ensure_vis(StreamCtx& C, size_t words) {
  if (C.vis_words >= words) return PANN_OK;
  if (C.d_vis) (void)hipFree(C.d_vis);
  C.d_vis = nullptr;
  C.vis_words = 0;
  HIP_TRY(dev_malloc(&C.d_vis, words * 8), "hipMalloc(visited lists)");
  C.vis_words = words;
  return PANN_OK;
}

// A launch's status word and work counter start at zero: one kernel clears both
// (a fill kernel, not a memset node, so a captured search replays it correctly).
__global__ void k_zero_launch(uint32_t* status, uint32_t* counter) {
  *status = 0u;
  *counter = 0u;
}

// A batch with per-query beams runs longest-first: a.qorder = the queries by
// descending beam (clamped to max_beam), sorted on the device.  The persistent
// grid takes queries in that order, so the long searches do not start last
// and set the batch's end.  Results do not depend on the order.  The sort
// clears the launch's status word and work counter (prepare_search was told
// to leave them); past its bins they are cleared on their own.
pann_status order_by_beam(StreamCtx& C, const int32_t* beam_pq, uint32_t nq, int64_t max_beam,
                          pann::SearchArgs& a, hipStream_t stream) {
  if (max_beam >= 8192) {  // past the sort's bins: the order given
    hipLaunchKernelGGL(k_zero_launch, dim3(1), dim3(1), 0, stream, a.status, a.work_counter);
    HIP_TRY(hipGetLastError(), "status and work counter reset");
    return PANN_OK;
  }
  const size_t need = nq;
  if (C.order_cap < need) {
    if (capturing(stream))
      return fail(PANN_ERR_ARG, "the query order scratch would grow under graph capture: run "
                                "the search once with at least this many queries before "
                                "capturing it");
    const size_t cap = std::max(need, 2 * C.order_cap);
    if (C.d_order) {
      HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
      (void)hipFree(C.d_order);
    }
    C.d_order = nullptr;
    C.order_cap = 0;
    HIP_TRY(dev_malloc(&C.d_order, cap * 4), "hipMalloc(query order)");
    C.order_cap = cap;
  }
  HIP_TRY(pann::launch_order_by_beam(beam_pq, nq, (uint32_t)max_beam, C.d_order, a.status,
                                     a.work_counter, stream),
          "query order launch");
  a.qorder = C.d_order;
  return PANN_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// pann_search_noalloc's memory: one file-scope device array per GPU, held
// until the process ends.  Bytes [0, 256) hold the persistent grid's work
// counter, [256, 512) the status word, [512, 768) the start points, and
// [1024, 8 MiB) the PANN_STORE_Q8 query codes (nq * dim bytes).  Host code
// never uses the array's name as an address: each device's address comes
// from hipGetSymbolAddress, taken once per device with that device current.
// ---------------------------------------------------------------------------
constexpr size_t kNoallocScratchBytes = 8u << 20;
constexpr size_t kNoallocCodesOffset = 1024;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

std::mutex g_noalloc_mu;              // the calls of this process share each device's array
void* g_noalloc_ptr[64] = {nullptr};  // its address per device

// The caller holds g_noalloc_mu and has made `device` current.
pann_status noalloc_scratch(int device, char** out) {
  if (device < 0 || device >= 64) return fail(PANN_ERR_ARG, "device ordinal out of range");
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_noalloc_ptr[device] = p;
  }
  *out = (char*)g_noalloc_ptr[device];
  return PANN_OK;
}

}  // namespace

extern "C" {

const char* pann_last_error(void) { return g_err.c_str(); }

pann_status pann_build_eval_counts(pann_index_t idx, uint64_t* search_evals,
                                   uint64_t* prune_evals, int32_t reset) {
  if (!idx) return fail(PANN_ERR_ARG, "null index");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  unsigned long long v[2];
  HIP_TRY(hipMemcpy(v, idx->d_evals, sizeof(v), hipMemcpyDeviceToHost), "hipMemcpy(counters)");
  if (search_evals) *search_evals = v[0];
  if (prune_evals) *prune_evals = v[1];
  if (reset) HIP_TRY(hipMemset(idx->d_evals, 0, sizeof(v)), "hipMemset(counters)");
  return PANN_OK;
}

pann_status pann_device_alloc_count(uint64_t* count) {
  if (!count) return fail(PANN_ERR_ARG, "null argument");
  *count = g_device_allocs.load(std::memory_order_relaxed);
  return PANN_OK;
}

pann_status pann_device_info(int32_t device, int32_t* num_cus, int64_t* lds_per_cu,
                             int64_t* hbm_bytes) {
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
  if (num_cus) *num_cus = p.multiProcessorCount;
  if (lds_per_cu) *lds_per_cu = (int64_t)p.maxSharedMemoryPerMultiProcessor;
  if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
  return PANN_OK;
}

pann_status pann_open(const pann_index_desc* desc, pann_index_t* out) {
  if (!desc || !out) return fail(PANN_ERR_ARG, "null argument");
  *out = nullptr;
  if (desc->n == 0 || desc->dim == 0 || desc->max_degree == 0)
    return fail(PANN_ERR_ARG, "empty index (n, dim and max_degree must be positive)");
  if (desc->dtype < PANN_DT_F32 || desc->dtype > PANN_DT_I8)
    return fail(PANN_ERR_ARG, "unknown dtype");
  if (desc->metric != PANN_METRIC_L2 && desc->metric != PANN_METRIC_MIPS)
    return fail(PANN_ERR_ARG, "unknown metric");
  if (desc->max_degree > 64)
    return fail(PANN_ERR_UNSUPPORTED, "max_degree > 64 is not supported by this build");
  if (!desc->points || !desc->graph) return fail(PANN_ERR_ARG, "null points or graph");
  const uint32_t n = desc->n, md = desc->max_degree;
  {
    pann_status cst = check_graph(n, md, desc->graph);
    if (cst != PANN_OK) return cst;
  }
  HIP_TRY(hipSetDevice(desc->device), "hipSetDevice");
  auto idx = new pann_index_s();
  idx->device = desc->device;
  idx->n = n;
  idx->dim = desc->dim;
  idx->max_degree = md;
  idx->dtype = desc->dtype;
  idx->metric = desc->metric;
  hipDeviceProp_t p;
  hipError_t e = hipGetDeviceProperties(&p, desc->device);
  if (e != hipSuccess) {
    delete idx;
    return hip_fail(e, "hipGetDeviceProperties");
  }
  idx->num_cus = p.multiProcessorCount;
  idx->lds_per_cu = (int64_t)p.maxSharedMemoryPerMultiProcessor;
  idx->max_lds_block = (int64_t)p.sharedMemPerBlockOptin > 0 ? (int64_t)p.sharedMemPerBlockOptin
                                                             : (int64_t)p.sharedMemPerBlock;
  if (idx->lds_per_cu <= 0) idx->lds_per_cu = 160 * 1024;
  const size_t eb = elem_bytes(desc->dtype);
  idx->row_bytes = (uint32_t)(desc->dim * eb);
  idx->pts_stride = round_up(idx->row_bytes, 128);
  idx->nbr_stride = (uint32_t)round_up(md, 16);
  auto cleanup = [&](pann_status st) {
    pann_close(idx);
    return st;
  };
  pann_status st;
  e = dev_malloc(&idx->d_pts, (size_t)n * idx->pts_stride);
  if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(points)"));
  {
    // rows padded with zeros on the host, uploaded in chunks with one plain
    // copy each (no separate device-side clear that a copy could overtake)
    const uint64_t rows = std::max<uint64_t>(1, (64ull << 20) / idx->pts_stride);
    std::vector<char> buf;
    for (uint64_t r0 = 0; r0 < n; r0 += rows) {
      uint64_t r1 = std::min<uint64_t>(n, r0 + rows);
      buf.assign((r1 - r0) * idx->pts_stride, 0);
      const char* src = (const char*)desc->points + r0 * idx->row_bytes;
      for (uint64_t i = r0; i < r1; i++)
        std::memcpy(&buf[(i - r0) * idx->pts_stride], src + (i - r0) * idx->row_bytes,
                    idx->row_bytes);
      e = hipMemcpy(idx->d_pts + r0 * idx->pts_stride, buf.data(), buf.size(),
                    hipMemcpyHostToDevice);
      if (e != hipSuccess) return cleanup(hip_fail(e, "hipMemcpy(points)"));
    }
  }
  e = dev_malloc(&idx->d_nbrs, (size_t)n * idx->nbr_stride * 4);
  if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(graph)"));
  st = upload_graph(n, md, idx->nbr_stride, desc->graph, idx->d_nbrs);
  if (st != PANN_OK) return cleanup(st);
  idx->levels.resize(1);
  idx->levels[0].d_nbrs = idx->d_nbrs;
  idx->levels[0].max_degree = md;
  idx->levels[0].nbr_stride = idx->nbr_stride;
  e = dev_malloc(&idx->d_dummy, 256);
  if (e != hipSuccess) return cleanup(hip_fail(e, "scratch allocation"));
  e = dev_malloc(&idx->d_evals, 2 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(idx->d_evals, 0, 2 * sizeof(unsigned long long));
  if (e != hipSuccess) return cleanup(hip_fail(e, "counter allocation"));
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return cleanup(hip_fail(e, "hipDeviceSynchronize"));
  (void)st;
  *out = idx;
  return PANN_OK;
}

namespace {

// PANN_STORE_Q8: per-row extrema on the device, the reference's parameter
// rules on the host, codes on the device.
pann_status build_q8(pann_index_t idx, uint64_t fstride) {
  const uint32_t n = idx->n;
  float *d_mins = nullptr, *d_maxs = nullptr;
  uint32_t* d_flag = nullptr;
  auto release = [&]() {
    for (void* p : {(void*)d_mins, (void*)d_maxs, (void*)d_flag})
      if (p) (void)hipFree(p);
  };
  hipError_t e = dev_malloc(&d_mins, (size_t)n * 4);
  if (e == hipSuccess) e = dev_malloc(&d_maxs, (size_t)n * 4);
  if (e == hipSuccess) e = dev_malloc(&d_flag, 256);
  if (e == hipSuccess)
    e = pann::launch_row_minmax((const float*)idx->d_pts, fstride, n, idx->dim, d_mins, d_maxs,
                                d_flag, nullptr);
  std::vector<float> mins(n), maxs(n);
  uint32_t notint = 0;
  if (e == hipSuccess) e = hipMemcpy(mins.data(), d_mins, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(maxs.data(), d_maxs, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&notint, d_flag, 4, hipMemcpyDeviceToHost);
  release();
  if (e != hipSuccess) return hip_fail(e, "Q8 extrema");
  float slope = 0, max_v = 0, scale = 0;
  int32_t offset = 0;
  if (idx->metric == PANN_METRIC_L2) {  // euclidian_point.h:219-243
    float min_val = *std::min_element(mins.begin(), mins.end());
    float max_val = *std::max_element(maxs.begin(), maxs.end());
    if (!notint) {
      if (max_val < 256) max_val = 255;
      min_val = 0;
    }
    slope = 255.0f / (max_val - min_val);
    offset = (int32_t)roundf(min_val * slope);
  } else {  // mips_point.h:415-470, trimmed
    const double cutoff = .0001;
    const size_t min_rank = (size_t)(cutoff * (double)n);
    const size_t max_rank = (size_t)((1.0 - cutoff) * (double)(n - 1));
    std::nth_element(mins.begin(), mins.begin() + min_rank, mins.end());
    std::nth_element(maxs.begin(), maxs.begin() + max_rank, maxs.end());
    const float min_val = mins[min_rank], max_val = maxs[max_rank];
    max_v = std::max(max_val, -min_val);
    scale = (float)(255 / 2) / max_v;
  }
  if (!idx->d_q8) {
    idx->q8_stride = round_up(idx->dim, 128);
    HIP_TRY(dev_malloc(&idx->d_q8, (size_t)n * idx->q8_stride), "hipMalloc(q8)");
    HIP_TRY(hipMemset(idx->d_q8, 0, (size_t)n * idx->q8_stride), "hipMemset(q8)");
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  }
  idx->q8_slope = slope;
  idx->q8_offset = offset;
  idx->q8_max = max_v;
  idx->q8_scale = scale;
  HIP_TRY(pann::launch_encode_q8((const float*)idx->d_pts, fstride, n, idx->dim,
                                 idx->metric == PANN_METRIC_L2 ? 0 : 1, slope, offset, max_v,
                                 scale, (uint8_t*)idx->d_q8, idx->q8_stride, nullptr),
          "Q8 encode");
  if (idx->metric == PANN_METRIC_MIPS) {
    if (!idx->d_q8_bound)
      HIP_TRY(dev_malloc(&idx->d_q8_bound, (size_t)n * 8), "hipMalloc(q8 bound)");
    HIP_TRY(pann::launch_q8_row_bound((const float*)idx->d_pts, fstride, n, idx->dim, scale,
                                      (const int8_t*)idx->d_q8, idx->q8_stride,
                                      idx->d_q8_bound, nullptr),
            "Q8 row bound kernel");
  }
  return PANN_OK;
}

// For PANN_STORE_Q8 the kernel reads query codes: encode the float queries
// into the index's scratch on `stream`.  For PANN_STORE_FP8_RESIDUAL it reads
// q . c for every centroid: one FP32 product on the matrix cores (k_qc_table)
// into the scratch.  Other stores read the queries as given.
pann_status stage_queries(pann_index_t idx, StreamCtx& C, int32_t store, const void* queries,
                          uint32_t nq, pann::SearchArgs& a, hipStream_t stream) {
  if (store == PANN_STORE_FP8_RESIDUAL) {
    const size_t need = (size_t)nq * idx->ncent;
    if (C.qtable_cap < need) {
      if (capturing(stream))
        return fail(PANN_ERR_ARG, "the centroid table would grow under graph capture: run the "
                                  "search once with at least this many queries before "
                                  "capturing it");
      if (C.d_qtable) {
        HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
        (void)hipFree(C.d_qtable);
      }
      C.d_qtable = nullptr;
      C.qtable_cap = 0;
      HIP_TRY(dev_malloc(&C.d_qtable, need * 4), "hipMalloc(centroid table)");
      C.qtable_cap = need;
    }
    // row-major T (nq x ncent) = Q (nq x dim) . C^T, C the centroid rows
    // (pts_stride bytes apart, zero padded like the points)
    HIP_TRY(pann::launch_qc_table((const float*)queries, idx->dim, nq, idx->dim, idx->d_cent,
                                  idx->pts_stride / 4, idx->ncent, C.d_qtable, stream),
            "centroid table launch");
    a.qtable = C.d_qtable;
    return PANN_OK;
  }
  if (store != PANN_STORE_Q8) return PANN_OK;
  const size_t need = (size_t)nq * idx->dim;
  if (C.qcodes_cap < need) {
    if (capturing(stream))
      return fail(PANN_ERR_ARG, "the query codes would grow under graph capture: run the "
                                "search once with at least this many queries before "
                                "capturing it");
    if (C.d_qcodes) {
      HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
      (void)hipFree(C.d_qcodes);
    }
    C.d_qcodes = nullptr;
    C.qcodes_cap = 0;
    HIP_TRY(dev_malloc(&C.d_qcodes, need), "hipMalloc(query codes)");
    C.qcodes_cap = need;
  }
  HIP_TRY(pann::launch_encode_q8((const float*)queries, idx->dim, nq, idx->dim,
                                 idx->metric == PANN_METRIC_L2 ? 0 : 1, idx->q8_slope,
                                 idx->q8_offset, idx->q8_max, idx->q8_scale, C.d_qcodes,
                                 idx->dim, stream),
          "query encode");
  a.queries = (const char*)C.d_qcodes;
  a.q_stride = idx->dim;
  return PANN_OK;
}

}  // namespace

pann_status pann_get_q8_params(pann_index_t idx, float* slope, int32_t* offset, float* max_val,
                               float* scale) {
  if (!idx) return fail(PANN_ERR_ARG, "null index");
  if (!idx->d_q8) return fail(PANN_ERR_ARG, "Q8 store not built");
  if (slope) *slope = idx->q8_slope;
  if (offset) *offset = idx->q8_offset;
  if (max_val) *max_val = idx->q8_max;
  if (scale) *scale = idx->q8_scale;
  return PANN_OK;
}

pann_status pann_build_store(pann_index_t idx, int32_t store) {
  if (!idx) return fail(PANN_ERR_ARG, "null index");
  if (store == PANN_STORE_FP8_RESIDUAL) return pann_build_fp8_residual(idx, 0);
  if (store != PANN_STORE_BF16 && store != PANN_STORE_FP8 && store != PANN_STORE_Q8)
    return fail(PANN_ERR_ARG,
                "store must be PANN_STORE_BF16, PANN_STORE_FP8, PANN_STORE_Q8 or "
                "PANN_STORE_FP8_RESIDUAL");
  if (idx->dtype != PANN_DT_F32)
    return fail(PANN_ERR_ARG, "reduced-precision stores need float points");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  const uint64_t fstride = idx->pts_stride / 4;
  if (store == PANN_STORE_Q8) {
    pann_status qst = build_q8(idx, fstride);
    if (qst != PANN_OK) return qst;
  } else if (store == PANN_STORE_BF16) {
    if (!idx->d_bf16) {
      idx->bf16_stride = round_up((uint64_t)idx->dim * 2, 128);
      HIP_TRY(dev_malloc(&idx->d_bf16, (size_t)idx->n * idx->bf16_stride), "hipMalloc(bf16)");
      HIP_TRY(hipMemset(idx->d_bf16, 0, (size_t)idx->n * idx->bf16_stride), "hipMemset(bf16)");
      HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }
    HIP_TRY(pann::launch_bf16_store((const float*)idx->d_pts, fstride, idx->n, idx->dim,
                                    (uint16_t*)idx->d_bf16, idx->bf16_stride / 2, nullptr),
            "bf16 store kernel");
    if (!idx->d_bf16_bound)
      HIP_TRY(dev_malloc(&idx->d_bf16_bound, (size_t)idx->n * 4), "hipMalloc(bf16 bound)");
    HIP_TRY(pann::launch_bf16_row_bound((const float*)idx->d_pts, fstride, idx->n, idx->dim,
                                        (const uint16_t*)idx->d_bf16, idx->bf16_stride / 2,
                                        idx->d_bf16_bound,
                                        idx->metric == PANN_METRIC_MIPS ? 1 : 0, nullptr),
            "bf16 row bound kernel");
  } else {
    if (!idx->d_fp8) {
      idx->fp8_stride = round_up((uint64_t)idx->dim, 128);
      HIP_TRY(dev_malloc(&idx->d_fp8, (size_t)idx->n * idx->fp8_stride), "hipMalloc(fp8)");
      HIP_TRY(hipMemset(idx->d_fp8, 0, (size_t)idx->n * idx->fp8_stride), "hipMemset(fp8)");
      HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
      HIP_TRY(dev_malloc(&idx->d_scales, idx->dim * 4), "hipMalloc(scales)");
      HIP_TRY(dev_malloc(&idx->d_absmax, idx->dim * 4), "hipMalloc(absmax)");
    }
    HIP_TRY(pann::launch_fp8_store((const float*)idx->d_pts, fstride, idx->n, idx->dim,
                                   idx->d_absmax, idx->d_scales, (uint8_t*)idx->d_fp8,
                                   idx->fp8_stride, nullptr),
            "fp8 store kernel");
    if (!idx->d_fp8_bound)
      HIP_TRY(dev_malloc(&idx->d_fp8_bound, (size_t)idx->n * 4), "hipMalloc(fp8 bound)");
    HIP_TRY(pann::launch_fp8_row_bound((const float*)idx->d_pts, fstride, idx->n, idx->dim,
                                       idx->d_scales, (const uint8_t*)idx->d_fp8,
                                       idx->fp8_stride, idx->d_fp8_bound,
                                       idx->metric == PANN_METRIC_MIPS ? 1 : 0, nullptr),
            "fp8 row bound kernel");
  }
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return PANN_OK;
}

pann_status pann_get_fp8_scales(pann_index_t idx, float* scales_host) {
  if (!idx || !scales_host) return fail(PANN_ERR_ARG, "null argument");
  if (!idx->d_scales) return fail(PANN_ERR_ARG, "FP8 store not built");
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  HIP_TRY(hipMemcpy(scales_host, idx->d_scales, idx->dim * 4, hipMemcpyDeviceToHost),
          "hipMemcpy(scales)");
  return PANN_OK;
}

namespace {

// Per-query start sets (device): query q starts from starts[q * stride ..
// + counts[q]) (counts NULL: one start each).  Counts are checked by the kernel.
struct StartSets {
  const uint32_t* starts = nullptr;
  const uint32_t* counts = nullptr;
  uint32_t stride = 0;
};

// Device memory a launch writes besides its outputs, when the caller brings
// it instead of the stream's context (pann_search_noalloc).
struct LaunchMem {
  uint32_t* counter;  // persistent grid's work counter
  uint32_t* status;   // status bits
  uint32_t* starts;   // start points (64 ids)
};

// Fills the launch block shared by every search entry point: graph `level`,
// host start points (starts, nstarts) or per-query device start sets (ss).
// With `fixed` the launch uses that memory and the stream's context is neither
// created nor used (C stays null): nothing is allocated.
pann_status prepare_search(pann_index_t idx, uint32_t level, const pann_qparams* qp,
                           const uint32_t* starts, uint32_t nstarts, const StartSets* ss,
                           const pann_search_opts* opts, uint32_t nq, pann::SearchArgs& a,
                           int& store_kind, bool& sequential, uint32_t& grid,
                           hipStream_t stream, StreamCtx*& C, uint32_t passes = 1,
                           const LaunchMem* fixed = nullptr, bool zero = true) {
  C = nullptr;
  if (!fixed) {
    pann_status cst = stream_ctx(idx, stream, &C);
    if (cst != PANN_OK) return cst;
  }
  if (level >= idx->levels.size() || !idx->levels[level].d_nbrs)
    return fail(PANN_ERR_ARG, "graph level " + std::to_string(level) + " is not set");
  const auto& G = idx->levels[level];
  const int64_t B = qp->beam_size;
  {
    pann_status qst = check_qparams(qp);
    if (qst != PANN_OK) return qst;
  }
  if (!ss) {
    if (nstarts == 0)  // beamSearch.h:51-53
      return fail(PANN_ERR_ARG, "beam search expects at least one start point");
    if (nstarts > B || nstarts > 64)  // :54-57
      return fail(PANN_ERR_ARG, "beam search has more starting points than beam size (or > 64)");
    std::unordered_set<uint32_t> seen;
    for (uint32_t i = 0; i < nstarts; i++) {
      if (starts[i] >= idx->n) return fail(PANN_ERR_ARG, "start point out of range");
      if (!seen.insert(starts[i]).second) return fail(PANN_ERR_ARG, "duplicate start points");
    }
  }
  pann_search_opts o{};
  if (opts) o = *opts;
  store_kind = pann::ST_F32;
  const char* pts = idx->d_pts;
  uint64_t stride = idx->pts_stride;
  uint32_t row_bytes = idx->row_bytes;
  sequential = false;
  if (o.store == PANN_STORE_NATIVE) {
    store_kind = idx->dtype == PANN_DT_F32 ? pann::ST_F32
                 : idx->dtype == PANN_DT_U8 ? pann::ST_U8 : pann::ST_I8;
    sequential = idx->dtype == PANN_DT_F32 && o.summation == PANN_SUM_INORDER;
  } else if (o.store == PANN_STORE_BF16) {
    if (!idx->d_bf16) return fail(PANN_ERR_ARG, "BF16 store not built (pann_build_store)");
    store_kind = pann::ST_BF16;
    pts = idx->d_bf16;
    stride = idx->bf16_stride;
    row_bytes = idx->dim * 2;
  } else if (o.store == PANN_STORE_Q8) {
    if (!idx->d_q8) return fail(PANN_ERR_ARG, "Q8 store not built (pann_build_store)");
    store_kind = idx->metric == PANN_METRIC_L2 ? pann::ST_U8 : pann::ST_I8;
    pts = idx->d_q8;
    stride = idx->q8_stride;
    row_bytes = idx->dim;
  } else if (o.store == PANN_STORE_FP8) {
    if (!idx->d_fp8) return fail(PANN_ERR_ARG, "FP8 store not built (pann_build_store)");
    store_kind = pann::ST_FP8;
    pts = idx->d_fp8;
    stride = idx->fp8_stride;
    row_bytes = idx->dim;
  } else if (o.store == PANN_STORE_FP8_RESIDUAL) {
    if (!idx->d_fp8r || idx->ncent == 0)  // ncent is set only once a build completed
      return fail(PANN_ERR_ARG, "FP8 residual store not built (pann_build_store)");
    store_kind = pann::ST_FP8R;
    pts = idx->d_fp8r;
    stride = idx->fp8r_stride;
    row_bytes = (uint32_t)idx->fp8r_stride;  // the centroid id sits in the last 4 bytes
  } else {
    return fail(PANN_ERR_ARG, "unknown store");
  }
  if (o.rerank && idx->dtype != PANN_DT_F32)
    return fail(PANN_ERR_ARG, "rerank needs float points");
  // the rerank keeps the first k of min(k * rerank_factor, |beam|) re-scored
  // entries (beamSearch.h:571-584): fewer than k would leave result slots
  // unwritten (the reference reads past its vector there)
  if (o.rerank && qp->rerank_factor < 1.0)
    return fail(PANN_ERR_ARG, "rerank needs rerank_factor >= 1");
  if (!sequential && row_bytes > 8 * 128)
    return fail(PANN_ERR_UNSUPPORTED, "rows wider than 1024 bytes need the in-order f32 path");

  std::memset(&a, 0, sizeof(a));
  a.pts = pts;
  a.pts_stride = stride;
  a.row_bytes = row_bytes;
  // 12-byte lane slices for rows of 1-byte elements that fill at most 96 bytes
  // of their line (the FP8 residual rows' last word names their centroid)
  a.lane_words = (!sequential && row_bytes <= 96 && store_kind != pann::ST_FP8R &&
                  (store_kind == pann::ST_FP8 || store_kind == pann::ST_U8 ||
                   store_kind == pann::ST_I8)) ? 3u : 4u;
  a.scales = store_kind == pann::ST_FP8R ? idx->d_rscales : idx->d_scales;
  a.ncent = store_kind == pann::ST_FP8R ? idx->ncent : 0u;
  a.full = o.rerank ? idx->d_pts : nullptr;
  a.full_stride = idx->pts_stride;
  a.row_bound = !o.rerank                   ? nullptr
                : store_kind == pann::ST_FP8 ? idx->d_fp8_bound
                : store_kind == pann::ST_BF16 ? idx->d_bf16_bound
                : o.store == PANN_STORE_Q8 && idx->metric == PANN_METRIC_MIPS ? idx->d_q8_bound
                                                                               : nullptr;
  a.q8_scale = idx->q8_scale;
  a.nbrs = G.d_nbrs;
  a.nbr_stride = G.nbr_stride;
  a.max_degree = G.max_degree;
  a.n = idx->n;
  a.dim = idx->dim;
  a.k = (int32_t)qp->k;
  a.beam = (int32_t)B;
  a.limit = qp->limit;
  a.degree_limit = (int32_t)std::min<int64_t>(qp->degree_limit, G.max_degree);
  a.batch_factor = qp->batch_factor;
  a.rerank_factor = qp->rerank_factor;
  a.status = fixed ? fixed->status : C->d_status;
  a.work_counter = fixed ? fixed->counter : C->d_counter;
  a.nstarts = ss ? 0u : nstarts;
  a.starts = fixed ? fixed->starts : C->d_starts;
  if (ss) {
    a.starts_pq = ss->starts;
    a.nstarts_pq = ss->counts;
    a.starts_stride = ss->stride;
  }

  // LDS layout
  const int64_t thr = (int64_t)std::ceil(qp->batch_factor * (double)B);
  const int64_t defer_cap = (qp->limit >= 2 * B) ? std::max<int64_t>(thr - 1, 0) : 0;
  a.fcap = (uint32_t)round_up((uint64_t)B, 16);
  a.ccap = (uint32_t)round_up((uint64_t)(defer_cap + a.degree_limit + 1), 16);
  a.ccap = std::max<uint32_t>(a.ccap, 64);  // also holds the start points
  a.exact_seen = o.exact_counts ? 1u : 0u;
  if (o.visits_per_round < 0 || o.visits_per_round > 3)
    return fail(PANN_ERR_ARG, "visits_per_round must be 0, 1, 2 or 3");
  // kernel lookahead: 0 none, 1 always, 2 when the merge rule predicts a deferral
  // (visits_per_round 0, the default, is 3: as fast or faster than 2 on every
  // workload measured)
  a.lookahead = o.visits_per_round == 1 ? 0u : (o.visits_per_round == 2 ? 1u : 2u);
  if (o.eval_cache_log2 != 0 && (o.eval_cache_log2 < 6 || o.eval_cache_log2 > 13))
    return fail(PANN_ERR_ARG, "eval_cache_log2 must be 0 or in [6, 13]");
  a.rc_log2 = o.exact_counts ? 0u : (uint32_t)o.eval_cache_log2;
  // frontier-id Bloom filter: 8 bits per frontier slot (2 of them set per
  // id), at most 8192 bits; its exact check scans the frontier 512 ids at a
  // time from registers.
  // Automatic: on for float32 rows wider than one 128-byte line, where a
  // gather it saves outweighs the filter test before every gather (measured
  // +3..7 %); the 1- and 2-line rows of the reduced-precision stores lose
  // more to that test than they save.
  if (o.frontier_filter < 0 || o.frontier_filter > 2)
    return fail(PANN_ERR_ARG, "frontier_filter must be 0, 1 or 2");
  if (o.waves_per_query != 0 && o.waves_per_query != 1 && o.waves_per_query != 2 &&
      o.waves_per_query != 4)
    return fail(PANN_ERR_ARG, "waves_per_query must be 0, 1, 2 or 4");
  if (o.query_order != 0 && o.query_order != 1)
    return fail(PANN_ERR_ARG, "query_order must be 0 or 1");
  a.wpq = o.waves_per_query > 1 ? (uint32_t)o.waves_per_query : 1u;
  if (a.wpq > 1 && (o.exact_counts || o.frontier_filter == 1))
    return fail(PANN_ERR_UNSUPPORTED, "waves_per_query > 1 runs the default seen set without "
                                      "the frontier filter (exact_counts = 0, frontier_filter "
                                      "0 or 2)");
  const bool fb_on =
      a.wpq == 1 && (o.frontier_filter == 1 ||
                     (o.frontier_filter == 0 && store_kind == pann::ST_F32 && a.row_bytes > 128));
  a.fb_log2 = (o.exact_counts || !fb_on)
                  ? 0u
                  : std::min<uint32_t>(std::max<uint32_t>(ceil_log2((uint64_t)8 * a.fcap), 9), 13);
  uint32_t hlog2;
  if (o.exact_counts) {
    // The reference's seen set (every id met): a small LDS table keeps more
    // waves per CU, ids past half its load go to the per-wave global table.
    hlog2 = (uint32_t)o.hash_slots_log2;
    if (hlog2 == 0) {
      double est = (double)B * 1.3 * (double)a.degree_limit * 0.45 * 2.0;
      hlog2 = std::min<uint32_t>(std::max<uint32_t>(ceil_log2((uint64_t)est), 10), 11u);
    }
    if (hlog2 < 8 || hlog2 > 15) return fail(PANN_ERR_ARG, "hash_slots_log2 must be in [8, 15]");
  } else {
    // The table only holds the pending candidates (at most ccap): half load.
    hlog2 = std::max<uint32_t>(ceil_log2((uint64_t)2 * a.ccap), 6);
  }
  a.hslots_log2 = hlog2;
  // the query row's own LDS region: for the in-order sum of rows past 8
  // lines, or when the row does not fit the two candidate buffers (the rerank's
  // staging place otherwise)
  a.q_own = (sequential && a.row_bytes > 8u * 128u) ||
            (uint64_t)idx->dim * 4 > 2ull * ((a.ccap * 8u + 15u) & ~15u);
  const uint32_t lds = pann::search_lds_bytes(a);
  if ((int64_t)lds > idx->max_lds_block || (int64_t)lds > idx->lds_per_cu)
    return fail(PANN_ERR_UNSUPPORTED, "beam too large for LDS (" + std::to_string(lds) + " bytes)");
  // queries per CU: waves_per_cu waves (16 by default) in workgroups of wpq
  // waves, and no more workgroups than the CU's LDS holds
  int wpc = o.waves_per_cu > 0 ? o.waves_per_cu : 16;
  int bpc = std::max<int>(1, wpc / (int)a.wpq);
  bpc = (int)std::min<int64_t>(bpc, idx->lds_per_cu / lds);
  if (bpc < 1) bpc = 1;
  grid = (uint32_t)(idx->num_cus * bpc);
  a.ghash = nullptr;
  a.ghash_log2 = 0;
  a.ghash_tag_base = 0;
  if (o.exact_counts && fixed)
    return fail(PANN_ERR_UNSUPPORTED, "exact_counts needs the overflow table, which is "
                                      "allocated: not available without allocation");
  if (o.exact_counts) {
    // per-wave global overflow table: twice the reference's initial hashset
    // size 2*(10+B)*degree_limit (beamSearch.h:66), at least 4x the LDS table
    uint32_t glog2 = std::max<uint32_t>(
        ceil_log2((uint64_t)2 * (10 + B) * a.degree_limit) + 1, hlog2 + 2);
    // grown after a search outgrew it, up to twice the ids there are (a
    // query sees each id once)
    const uint32_t gmax = std::max<uint32_t>(glog2, ceil_log2(2 * ((uint64_t)idx->n + 64)) + 1);
    glog2 = std::min<uint32_t>(glog2 + idx->ghash_grow.load(), std::min<uint32_t>(gmax, 31u));
    idx->ghash_log2_last.store(glog2);
    a.ghash_log2 = glog2;
    // every pass of every query gets its own tag (the HNSW descent runs
    // passes * nq searches in one launch)
    pann_status st = ensure_ghash(*C, (size_t)grid << glog2, nq * passes);
    if (st != PANN_OK) return st;
    a.ghash = C->d_ghash;
    a.ghash_tag_base = C->ghash_tag;
    C->ghash_tag += nq * passes;
  }
  // start points go up only when they change, so a repeated search issues no
  // host-to-device copy (and can be captured into a graph)
  const bool cap = capturing(stream);
  if (cap && o.exact_counts)
    return fail(PANN_ERR_ARG, "exact_counts searches cannot be captured into a graph");
  if (fixed && !ss) {  // no per-stream record of what the memory holds: copy each time
    HIP_TRY(hipMemcpyAsync(fixed->starts, starts, nstarts * 4, hipMemcpyHostToDevice, stream),
            "hipMemcpyAsync(starts)");
  } else if (!ss && (C->starts_host.size() != nstarts ||
                     !std::equal(starts, starts + nstarts, C->starts_host.begin()))) {
    if (cap)
      return fail(PANN_ERR_ARG, "start points changed under graph capture: run the search "
                                "once with these start points before capturing it");
    HIP_TRY(hipMemcpyAsync(C->d_starts, starts, nstarts * 4, hipMemcpyHostToDevice, stream),
            "hipMemcpyAsync(starts)");
    HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    C->starts_host.assign(starts, starts + nstarts);
  }
  if (zero) {  // else the caller's order_by_beam clears them
    hipLaunchKernelGGL(k_zero_launch, dim3(1), dim3(1), 0, stream, a.status, a.work_counter);
    HIP_TRY(hipGetLastError(), "status and work counter reset");
  }
  return PANN_OK;
}

// set when a synchronous status check finds the exact seen set's table
// outgrown (grow_retry below grows it and runs the search again)
thread_local bool tl_seen_table_full = false;

// allow_short: a frontier shorter than k is a result, not an error (HNSW
// layers return whatever frontier they reach)
pann_status check_status_word(const uint32_t* d_status, hipStream_t stream,
                              bool allow_short = false) {
  HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
  uint32_t st = 0;
  HIP_TRY(hipMemcpy(&st, d_status, 4, hipMemcpyDeviceToHost), "hipMemcpy(status)");
  if (st & pann::PANN_ST_BAD_PARAMS)  // beamSearch.h:612-616 (k > beam) per query
    return fail(PANN_ERR_ARG, "a per-query k or beam is invalid (k > beam, beam > qp->beam_size, "
                              "or k > out_stride)");
  if ((st & pann::PANN_ST_SHORT_BEAM) && !allow_short)  // beamSearch.h:557-562
    return fail(PANN_ERR_SHORT_BEAM, "beam search returned fewer than k elements");
  if (st & pann::PANN_ST_HASH_CAPACITY) {
    tl_seen_table_full = true;
    return fail(PANN_ERR_CAPACITY, "a query outgrew the exact seen set's table (an async "
                                   "search cannot grow it: run it synchronously once)");
  }
  if (st & pann::PANN_ST_VISITED_OVERFLOW)
    return fail(PANN_ERR_CAPACITY, "a visited list outgrew vis_cap");
  return PANN_OK;
}

pann_status check_status(StreamCtx& C, hipStream_t stream, bool allow_short = false) {
  return check_status_word(C.d_status, stream, allow_short);
}

// Runs a synchronous search; when a query outgrew the exact seen set's
// global table (the reference's hashset grows instead, hashset.h:25-38),
// grows the table 4x and runs the whole search again -- its results do not
// depend on the table's size -- until it fits or the table holds every id.
pann_status grow_retry(pann_index_t idx, const std::function<pann_status()>& run) {
  for (;;) {
    tl_seen_table_full = false;
    const pann_status st = run();
    if (st != PANN_ERR_CAPACITY || !tl_seen_table_full || !idx) return st;
    // the launch's table had 2^used slots per wave; at the size that holds
    // every id of the index no query can outgrow it
    const uint32_t used = idx->ghash_log2_last.load();
    const uint32_t gmax = std::min<uint32_t>(ceil_log2(2 * ((uint64_t)idx->n + 64)) + 1, 31u);
    if (used >= gmax) return st;
    idx->ghash_grow.fetch_add(2);
  }
}

}  // namespace

namespace {

// What the frontier and visited-list entry points add to pann_search.
struct SearchExtras {
  uint32_t* out_nfront = nullptr;  // device: frontier size per query
  bool allow_short = false;        // a frontier shorter than k is a result, not an error
  uint32_t vis_cap = 0;            // > 0: write each query's visited list, sorted by
  uint32_t* vis_ids = nullptr;     //   (dist, id), to vis_ids / vis_dists (device,
  float* vis_dists = nullptr;      //   nq x vis_cap; counts in out_visited)
};

pann_status search_impl(pann_index_t idx, const void* queries, uint32_t nq,
                        const pann_qparams* qp, const int32_t* k_pq, const int32_t* beam_pq,
                        const uint32_t* starts, uint32_t nstarts, const pann_search_opts* opts,
                        uint32_t* out_ids, float* out_dists, uint32_t out_stride,
                        uint32_t* out_visited, uint32_t* out_cmps, const SearchExtras& ex,
                        void* stream_) {
  if (!idx || !qp || !starts) return fail(PANN_ERR_ARG, "null argument");
  {
    pann_status qst = check_qparams(qp);
    if (qst != PANN_OK) return qst;
  }
  if (nq == 0) return PANN_OK;
  if (!queries || !out_ids || !out_dists) return fail(PANN_ERR_ARG, "null buffer");
  if (out_stride < (uint32_t)std::max<int64_t>(qp->k, 0))
    return fail(PANN_ERR_ARG, "out_stride smaller than k");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  hipStream_t stream = (hipStream_t)stream_;
  pann::SearchArgs a;
  int store_kind;
  bool sequential;
  uint32_t grid;
  StreamCtx* C = nullptr;
  const bool will_order = beam_pq && !(opts && opts->query_order);  // order_by_beam below
  pann_status st = prepare_search(idx, 0, qp, starts, nstarts, nullptr, opts, nq, a, store_kind,
                                  sequential, grid, stream, C, 1, nullptr, !will_order);
  if (st != PANN_OK) return st;
  const bool f32_queries = idx->dtype == PANN_DT_F32;
  a.queries = (const char*)queries;
  a.q_stride = (uint64_t)idx->dim * (f32_queries ? 4 : 1);
  a.fqueries = (const char*)queries;
  a.fq_stride = (uint64_t)idx->dim * 4;
  a.nq = nq;
  a.k_pq = k_pq;
  a.beam_pq = beam_pq;
  a.out_ids = out_ids;
  a.out_dists = out_dists;
  a.out_stride = out_stride;
  a.out_visited = out_visited;
  a.out_cmps = out_cmps;
  a.out_nfront = ex.out_nfront;
  a.short_ok = ex.allow_short ? 1u : 0u;
  if (ex.vis_cap) {
    if (!out_visited || !ex.vis_ids || !ex.vis_dists)
      return fail(PANN_ERR_ARG, "visited lists need out_visited, ids and distances");
    st = ensure_vis(*C, (size_t)nq * ex.vis_cap);
    if (st != PANN_OK) return st;
    a.vis_out = C->d_vis;
    a.vis_cap = ex.vis_cap;
  }
  st = stage_queries(idx, *C, opts ? opts->store : PANN_STORE_NATIVE, queries, nq, a, stream);
  if (st != PANN_OK) return st;
  if (will_order) {
    st = order_by_beam(*C, beam_pq, nq, qp->beam_size, a, stream);
    if (st != PANN_OK) return st;
  }
  grid = std::min<uint32_t>(grid, nq);
  const bool cap = capturing(stream);  // captured: no events, status read after replay
  if (cap && !(opts && opts->async))
    return fail(PANN_ERR_ARG, "a search under graph capture must be async");
  if (!cap) HIP_TRY(hipEventRecord(C->ev0, stream), "hipEventRecord");
  HIP_TRY(pann::launch_beam_search(a, store_kind, idx->metric, sequential, grid, stream),
          "beam search launch");
  if (!cap) {
    HIP_TRY(hipEventRecord(C->ev1, stream), "hipEventRecord");
    idx->last_timed = C;
  }
  if (ex.vis_cap)
    HIP_TRY(pann::launch_sort_visited(C->d_vis, nq, ex.vis_cap, out_visited, ex.vis_ids,
                                      ex.vis_dists, stream),
            "visited sort launch");
  if (opts && opts->async) return PANN_OK;
  return check_status(*C, stream, ex.allow_short);
}

}  // namespace

pann_status pann_search(pann_index_t idx, const void* queries, uint32_t nq,
                        const pann_qparams* qp, const int32_t* k_pq, const int32_t* beam_pq,
                        const uint32_t* starts, uint32_t nstarts, const pann_search_opts* opts,
                        uint32_t* out_ids, float* out_dists, uint32_t out_stride,
                        uint32_t* out_visited, uint32_t* out_cmps, void* stream) {
  return grow_retry(idx, [&] {
    return search_impl(idx, queries, nq, qp, k_pq, beam_pq, starts, nstarts, opts, out_ids,
                       out_dists, out_stride, out_visited, out_cmps, SearchExtras(), stream);
  });
}

pann_status pann_search_noalloc(pann_index_t idx, const void* queries, uint32_t nq,
                                const pann_qparams* qp, const int32_t* k_pq,
                                const int32_t* beam_pq, const uint32_t* starts, uint32_t nstarts,
                                const pann_search_opts* opts, uint32_t* out_ids, float* out_dists,
                                uint32_t out_stride, uint32_t* out_visited, uint32_t* out_cmps,
                                void* stream_) {
  if (!idx || !qp || !starts) return fail(PANN_ERR_ARG, "null argument");
  {
    pann_status qst = check_qparams(qp);
    if (qst != PANN_OK) return qst;
  }
  if (nq == 0) return PANN_OK;
  if (!queries || !out_ids || !out_dists) return fail(PANN_ERR_ARG, "null buffer");
  if (out_stride < (uint32_t)std::max<int64_t>(qp->k, 0))
    return fail(PANN_ERR_ARG, "out_stride smaller than k");
  pann_search_opts o{};
  if (opts) o = *opts;
  if (o.async) return fail(PANN_ERR_ARG, "pann_search_noalloc is synchronous (async = 0)");
  if (o.store == PANN_STORE_FP8_RESIDUAL)
    return fail(PANN_ERR_UNSUPPORTED, "the FP8 residual store's per-query table is allocated: "
                                      "not available without allocation");
  if (o.store == PANN_STORE_Q8 &&
      (uint64_t)nq * idx->dim > kNoallocScratchBytes - kNoallocCodesOffset)
    return fail(PANN_ERR_CAPACITY,
                "PANN_STORE_Q8 without allocation: nq * dim = " +
                    std::to_string((uint64_t)nq * idx->dim) + " exceeds " +
                    std::to_string(kNoallocScratchBytes - kNoallocCodesOffset) + " bytes");
  std::lock_guard<std::mutex> gl(g_noalloc_mu);
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  char* p = nullptr;
  {
    pann_status st = noalloc_scratch(idx->device, &p);
    if (st != PANN_OK) return st;
  }
  const LaunchMem lm{(uint32_t*)p, (uint32_t*)(p + 256), (uint32_t*)(p + 512)};
  hipStream_t stream = (hipStream_t)stream_;
  pann::SearchArgs a;
  int store_kind;
  bool sequential;
  uint32_t grid;
  StreamCtx* C = nullptr;
  pann_status st = prepare_search(idx, 0, qp, starts, nstarts, nullptr, &o, nq, a, store_kind,
                                  sequential, grid, stream, C, 1, &lm);
  if (st != PANN_OK) return st;
  const bool f32_queries = idx->dtype == PANN_DT_F32;
  a.queries = (const char*)queries;
  a.q_stride = (uint64_t)idx->dim * (f32_queries ? 4 : 1);
  a.fqueries = (const char*)queries;
  a.fq_stride = (uint64_t)idx->dim * 4;
  a.nq = nq;
  a.k_pq = k_pq;
  a.beam_pq = beam_pq;
  a.out_ids = out_ids;
  a.out_dists = out_dists;
  a.out_stride = out_stride;
  a.out_visited = out_visited;
  a.out_cmps = out_cmps;
  if (o.store == PANN_STORE_Q8) {  // the query codes go to the array past its first KiB
    uint8_t* codes = (uint8_t*)(p + kNoallocCodesOffset);
    HIP_TRY(pann::launch_encode_q8((const float*)queries, idx->dim, nq, idx->dim,
                                   idx->metric == PANN_METRIC_L2 ? 0 : 1, idx->q8_slope,
                                   idx->q8_offset, idx->q8_max, idx->q8_scale, codes, idx->dim,
                                   stream),
            "query encode");
    a.queries = (const char*)codes;
    a.q_stride = idx->dim;
  }
  grid = std::min<uint32_t>(grid, nq);
  HIP_TRY(pann::launch_beam_search(a, store_kind, idx->metric, sequential, grid, stream),
          "beam search launch");
  return check_status_word(lm.status, stream);
}

pann_status pann_search_status(pann_index_t idx, void* stream) {
  if (!idx) return fail(PANN_ERR_ARG, "null index");
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  StreamCtx* C = nullptr;
  {
    std::lock_guard<std::mutex> lk(idx->mu);
    pann_status st = stream_ctx(idx, (hipStream_t)stream, &C);
    if (st != PANN_OK) return st;
  }
  return check_status(*C, (hipStream_t)stream);
}

pann_status pann_last_kernel_ms(pann_index_t idx, float* ms) {
  if (!idx || !ms) return fail(PANN_ERR_ARG, "null argument");
  StreamCtx* C = idx->last_timed;
  if (!C) return fail(PANN_ERR_ARG, "no search has run");
  HIP_TRY(hipEventSynchronize(C->ev1), "hipEventSynchronize");
  HIP_TRY(hipEventElapsedTime(ms, C->ev0, C->ev1), "hipEventElapsedTime");
  return PANN_OK;
}

namespace {

// Carves the staging buffer of the default stream into consecutive regions
// (256-byte aligned), growing it first.  The caller holds idx->host_mu.
struct Stage {
  size_t off = 0;
  std::vector<std::pair<void**, size_t>> regions;
  void add(void** p, size_t bytes) {
    regions.push_back({p, off});
    off += round_up(std::max<size_t>(bytes, 1), 256);
  }
  pann_status commit(pann_index_t idx) {
    StreamCtx* C = nullptr;
    {
      std::lock_guard<std::mutex> lk(idx->mu);
      pann_status st = stream_ctx(idx, nullptr, &C);
      if (st != PANN_OK) return st;
      st = ensure_stage(*C, off);
      if (st != PANN_OK) return st;
    }
    for (auto& r : regions) *r.first = C->d_stage + r.second;
    return PANN_OK;
  }
};

// pann_beam_search_host and pann_beam_search_visited_host: the whole final
// frontier, and optionally the sorted visited list.
pann_status beam_search_host(pann_index_t idx, const void* queries, uint32_t nq,
                             const pann_qparams* qp, const uint32_t* starts, uint32_t nstarts,
                             const pann_search_opts* opts, uint32_t* out_ids, float* out_dists,
                             uint32_t* out_count, uint32_t* out_visited, uint32_t vis_cap,
                             uint32_t* out_vis_ids, float* out_vis_dists, uint32_t* out_cmps) {
  if (!idx || !qp || !queries || !out_ids || !out_dists || !out_count)
    return fail(PANN_ERR_ARG, "null argument");
  const bool want_vis = out_vis_ids || out_vis_dists;
  if (want_vis && !(out_vis_ids && out_vis_dists && out_visited))
    return fail(PANN_ERR_ARG, "the visited outputs come together (ids, distances, counts)");
  if (want_vis && (vis_cap == 0 || vis_cap > 4096))
    return fail(PANN_ERR_ARG, "vis_cap must be in [1, 4096]");
  if (nq == 0) return PANN_OK;
  if (qp->beam_size < 1 || qp->beam_size > 4096)
    return fail(PANN_ERR_ARG, "beam_size must be in [1, 4096]");
  std::lock_guard<std::mutex> hl(idx->host_mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  const size_t qbytes = (size_t)nq * idx->dim * (idx->dtype == PANN_DT_F32 ? 4 : 1);
  const size_t B = (size_t)qp->beam_size;
  pann_qparams q2 = *qp;
  q2.k = qp->beam_size;  // the whole frontier comes back
  void* dq = nullptr;
  uint32_t *di = nullptr, *dn = nullptr, *dv = nullptr, *dc = nullptr, *dvi = nullptr;
  float *dd = nullptr, *dvd = nullptr;
  Stage sg;
  sg.add((void**)&dq, qbytes);
  sg.add((void**)&di, nq * B * 4);
  sg.add((void**)&dd, nq * B * 4);
  sg.add((void**)&dn, nq * 4);
  sg.add((void**)&dv, nq * 4);
  sg.add((void**)&dc, nq * 4);
  if (want_vis) {
    sg.add((void**)&dvi, (size_t)nq * vis_cap * 4);
    sg.add((void**)&dvd, (size_t)nq * vis_cap * 4);
  }
  pann_status st = sg.commit(idx);
  if (st != PANN_OK) return st;
  HIP_TRY(hipMemcpy(dq, queries, qbytes, hipMemcpyHostToDevice), "beam_search_host staging");
  pann_search_opts o{};
  if (opts) o = *opts;
  o.async = 0;
  o.rerank = 0;
  SearchExtras ex;
  ex.out_nfront = dn;
  ex.allow_short = true;
  ex.vis_cap = want_vis ? vis_cap : 0u;
  ex.vis_ids = dvi;
  ex.vis_dists = dvd;
  st = grow_retry(idx, [&] {
    return search_impl(idx, dq, nq, &q2, nullptr, nullptr, starts, nstarts, &o, di, dd,
                       (uint32_t)B, dv, dc, ex, nullptr);
  });
  if (st != PANN_OK) return st;
  hipError_t e = hipMemcpy(out_ids, di, nq * B * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(out_dists, dd, nq * B * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(out_count, dn, nq * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && out_cmps) e = hipMemcpy(out_cmps, dc, nq * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && out_visited)
    e = hipMemcpy(out_visited, dv, nq * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && want_vis) {
    e = hipMemcpy(out_vis_ids, dvi, (size_t)nq * vis_cap * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(out_vis_dists, dvd, (size_t)nq * vis_cap * 4, hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) return hip_fail(e, "beam_search_host copy-out");
  return PANN_OK;
}

}  // namespace

pann_status pann_search_host(pann_index_t idx, const void* queries, uint32_t nq,
                             const pann_qparams* qp, const uint32_t* starts, uint32_t nstarts,
                             const pann_search_opts* opts, uint32_t* out_ids, float* out_dists,
                             uint32_t* out_visited, uint32_t* out_cmps) {
  if (!idx || !qp || !queries || !out_ids || !out_dists) return fail(PANN_ERR_ARG, "null argument");
  if (qp->k < 1) return fail(PANN_ERR_ARG, "k must be >= 1");
  {
    pann_status qst = check_qparams(qp);
    if (qst != PANN_OK) return qst;
  }
  if (nq == 0) return PANN_OK;
  std::lock_guard<std::mutex> hl(idx->host_mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  const size_t qbytes = (size_t)nq * idx->dim * (idx->dtype == PANN_DT_F32 ? 4 : 1);
  const size_t k = (size_t)qp->k;
  void* dq = nullptr;
  uint32_t *di = nullptr, *dv = nullptr, *dc = nullptr;
  float* dd = nullptr;
  Stage sg;
  sg.add((void**)&dq, qbytes);
  sg.add((void**)&di, nq * k * 4);
  sg.add((void**)&dd, nq * k * 4);
  sg.add((void**)&dv, nq * 4);
  sg.add((void**)&dc, nq * 4);
  pann_status st = sg.commit(idx);
  if (st != PANN_OK) return st;
  HIP_TRY(hipMemcpy(dq, queries, qbytes, hipMemcpyHostToDevice), "search_host staging");
  pann_search_opts o{};
  if (opts) o = *opts;
  o.async = 0;
  st = pann_search(idx, dq, nq, qp, nullptr, nullptr, starts, nstarts, &o, di, dd, (uint32_t)k,
                   dv, dc, nullptr);
  if (st != PANN_OK) return st;
  hipError_t e = hipMemcpy(out_ids, di, nq * k * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(out_dists, dd, nq * k * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && out_visited) e = hipMemcpy(out_visited, dv, nq * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && out_cmps) e = hipMemcpy(out_cmps, dc, nq * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "search_host copy-out");
  return PANN_OK;
}

pann_status pann_beam_search_host(pann_index_t idx, const void* queries, uint32_t nq,
                                  const pann_qparams* qp, const uint32_t* starts,
                                  uint32_t nstarts, const pann_search_opts* opts,
                                  uint32_t* out_ids, float* out_dists, uint32_t* out_count,
                                  uint32_t* out_visited, uint32_t* out_cmps) {
  return beam_search_host(idx, queries, nq, qp, starts, nstarts, opts, out_ids, out_dists,
                          out_count, out_visited, 0, nullptr, nullptr, out_cmps);
}

pann_status pann_beam_search_visited_host(pann_index_t idx, const void* queries, uint32_t nq,
                                          const pann_qparams* qp, const uint32_t* starts,
                                          uint32_t nstarts, const pann_search_opts* opts,
                                          uint32_t* out_ids, float* out_dists,
                                          uint32_t* out_count, uint32_t vis_cap,
                                          uint32_t* out_vis_ids, float* out_vis_dists,
                                          uint32_t* out_vis_count, uint32_t* out_cmps) {
  if (!out_vis_ids || !out_vis_dists || !out_vis_count)
    return fail(PANN_ERR_ARG, "null visited output");
  return beam_search_host(idx, queries, nq, qp, starts, nstarts, opts, out_ids, out_dists,
                          out_count, out_vis_count, vis_cap, out_vis_ids, out_vis_dists,
                          out_cmps);
}

pann_status pann_search_visited(pann_index_t idx, const uint32_t* self_ids, uint32_t nq,
                                uint32_t beam, double batch_factor, uint32_t start,
                                uint32_t vis_cap, uint32_t* vis_ids, float* vis_dists,
                                uint32_t* vis_count, void* stream_) {
  if (!idx || !self_ids || !vis_ids || !vis_dists || !vis_count)
    return fail(PANN_ERR_ARG, "null argument");
  if (nq == 0) return PANN_OK;
  if (vis_cap == 0 || vis_cap > 4096) return fail(PANN_ERR_ARG, "vis_cap must be in [1, 4096]");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  hipStream_t stream = (hipStream_t)stream_;
  // QueryParams QP(0, L, n, max_degree, rerank, batch_factor) (vamana/index.h:251-252)
  pann_qparams qp{0, (int64_t)beam, (int64_t)idx->n, (int64_t)idx->max_degree, 100.0,
                  batch_factor};
  pann_search_opts o{};
  o.store = PANN_STORE_NATIVE;
  o.summation = PANN_SUM_INORDER;
  pann::SearchArgs a;
  int store_kind;
  bool sequential;
  uint32_t grid;
  StreamCtx* C = nullptr;
  pann_status st = prepare_search(idx, 0, &qp, &start, 1, nullptr, &o, nq, a, store_kind,
                                  sequential, grid, stream, C);
  if (st != PANN_OK) return st;
  st = ensure_vis(*C, (size_t)nq * vis_cap);
  if (st != PANN_OK) return st;
  if (C->cmps_cap < nq) {  // per-query evaluations, summed into the build counter
    const size_t cap = std::max<size_t>((size_t)nq, 2 * C->cmps_cap);
    if (C->d_cmps) (void)hipFree(C->d_cmps);
    C->d_cmps = nullptr;
    C->cmps_cap = 0;
    HIP_TRY(dev_malloc(&C->d_cmps, cap * 4), "hipMalloc(evaluation counts)");
    C->cmps_cap = cap;
  }
  a.queries = idx->d_pts;  // each query is the point's own row
  a.q_stride = idx->pts_stride;
  a.fqueries = idx->d_pts;
  a.fq_stride = idx->pts_stride;
  a.nq = nq;
  a.self_ids = self_ids;
  a.vis_out = C->d_vis;
  a.vis_cap = vis_cap;
  a.out_ids = idx->d_dummy;
  a.out_dists = (float*)idx->d_dummy;
  a.out_stride = 0;
  a.out_visited = vis_count;
  a.out_cmps = C->d_cmps;
  a.accumulate = 0u;
  // with self_ids set the kernel reads query q from row self_ids[q]
  grid = std::min<uint32_t>(grid, nq);
  HIP_TRY(pann::launch_beam_search(a, store_kind, idx->metric, sequential, grid, stream),
          "build-mode search launch");
  HIP_TRY(pann::launch_sum_u32(C->d_cmps, nq, idx->d_evals, stream), "evaluation count launch");
  HIP_TRY(pann::launch_sort_visited(C->d_vis, nq, vis_cap, vis_count, vis_ids, vis_dists, stream),
          "visited sort launch");
  return check_status(*C, stream);
}

namespace {

pann_status run_prune(pann_index_t idx, uint32_t level, const uint32_t* pts, uint32_t np,
                      const uint32_t* cand_ids, const float* cand_dists,
                      const uint32_t* cand_count, uint32_t cand_stride, const uint32_t* offsets,
                      uint32_t max_count, double alpha, uint32_t R, int32_t add,
                      uint32_t* out_ids, uint32_t* out_count, hipStream_t stream) {
  if (level >= idx->levels.size() || !idx->levels[level].d_nbrs)
    return fail(PANN_ERR_ARG, "graph level " + std::to_string(level) + " is not set");
  const auto& G = idx->levels[level];
  if (R == 0 || R > G.max_degree) return fail(PANN_ERR_ARG, "R must be in [1, max_degree]");
  if (add < 0 || add > 5) return fail(PANN_ERR_ARG, "unknown prune mode");
  pann::PruneArgs a;
  std::memset(&a, 0, sizeof(a));
  a.pts = idx->d_pts;
  a.stride = idx->pts_stride;
  a.dim = idx->dim;
  a.nbrs = G.d_nbrs;
  a.nbr_stride = G.nbr_stride;
  a.max_degree = G.max_degree;
  a.points = pts;
  a.np = np;
  a.cand_ids = cand_ids;
  a.cand_dists = cand_dists;
  a.cand_count = cand_count;
  a.cand_stride = cand_stride;
  a.cand_offsets = offsets;
  a.alpha = alpha;
  a.R = R;
  a.add = add;
  a.out_ids = out_ids;
  a.out_count = out_count;
  a.evals = idx->d_evals + 1;
  uint64_t cap = 2;
  while (cap < (uint64_t)max_count + G.max_degree) cap <<= 1;
  if (cap > (1u << 24)) return fail(PANN_ERR_UNSUPPORTED, "candidate set too large");
  a.sort_cap = (uint32_t)cap;
  const int dk = idx->dtype == PANN_DT_F32 ? pann::ST_F32
                 : idx->dtype == PANN_DT_U8 ? pann::ST_U8 : pann::ST_I8;
  const uint32_t lds = pann::prune_lds_bytes(a.sort_cap, idx->dim, dk);
  const bool keys_global = (int64_t)lds > std::min<int64_t>(idx->max_lds_block, 64 * 1024);
  uint32_t grid = std::min<uint32_t>(np, 1u << 20);
  if (keys_global) {
    grid = std::min<uint32_t>(np, (uint32_t)idx->num_cus * 4);
    size_t need = (size_t)grid * cap;
    if (idx->pscratch_entries < need) {
      if (idx->d_pkeys) (void)hipFree(idx->d_pkeys);
      if (idx->d_pflags) (void)hipFree(idx->d_pflags);
      idx->d_pkeys = nullptr;
      idx->d_pflags = nullptr;
      idx->pscratch_entries = 0;
      HIP_TRY(dev_malloc(&idx->d_pkeys, need * 8), "hipMalloc(prune scratch)");
      HIP_TRY(dev_malloc(&idx->d_pflags, need), "hipMalloc(prune scratch)");
      idx->pscratch_entries = need;
    }
    a.gkeys = idx->d_pkeys;
    a.gflags = idx->d_pflags;
  }
  HIP_TRY(pann::launch_robust_prune(a, dk, idx->metric, keys_global, grid, stream),
          "robust prune launch");
  HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
  return PANN_OK;
}

}  // namespace

pann_status pann_robust_prune_batch(pann_index_t idx, const uint32_t* pts, uint32_t np,
                                    const uint32_t* cand_ids, const float* cand_dists,
                                    const uint32_t* cand_count, uint32_t cand_stride,
                                    double alpha, uint32_t R, int32_t add, uint32_t* out_ids,
                                    uint32_t* out_count, void* stream) {
  if (!idx || !pts || !cand_ids || !cand_count || !out_ids || !out_count)
    return fail(PANN_ERR_ARG, "null argument");
  if (np == 0) return PANN_OK;
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  if (add > 3) return fail(PANN_ERR_ARG, "add must be 0, 1, 2 or 3");
  return run_prune(idx, 0, pts, np, cand_ids, cand_dists, cand_count, cand_stride, nullptr,
                   cand_stride, alpha, R, add, out_ids, out_count, (hipStream_t)stream);
}

pann_status pann_robust_prune_groups(pann_index_t idx, const uint32_t* pts, uint32_t np,
                                     const uint32_t* offsets, const uint32_t* cand_ids,
                                     uint32_t max_count, double alpha, uint32_t R, int32_t add,
                                     uint32_t* out_ids, uint32_t* out_count, void* stream) {
  if (!idx || !pts || !offsets || !cand_ids || !out_ids || !out_count)
    return fail(PANN_ERR_ARG, "null argument");
  if (np == 0) return PANN_OK;
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  if (add > 3) return fail(PANN_ERR_ARG, "add must be 0, 1, 2 or 3");
  return run_prune(idx, 0, pts, np, cand_ids, nullptr, nullptr, 0, offsets, max_count, alpha, R,
                   add, out_ids, out_count, (hipStream_t)stream);
}

namespace {

pann_status update_level(pann_index_t idx, uint32_t level, const uint32_t* pts, uint32_t np,
                         const uint32_t* lists, const uint32_t* counts, uint32_t R,
                         hipStream_t stream) {
  if (level >= idx->levels.size() || !idx->levels[level].d_nbrs)
    return fail(PANN_ERR_ARG, "graph level " + std::to_string(level) + " is not set");
  const auto& G = idx->levels[level];
  if (R == 0 || R > G.max_degree) return fail(PANN_ERR_ARG, "R must be in [1, max_degree]");
  HIP_TRY(pann::launch_update_neighbors(G.d_nbrs, G.nbr_stride, G.max_degree, pts, np, lists,
                                        counts, R, stream),
          "update neighbours launch");
  HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
  return PANN_OK;
}

pann_status download_level(pann_index_t idx, uint32_t level, uint32_t* graph_host) {
  if (level >= idx->levels.size() || !idx->levels[level].d_nbrs)
    return fail(PANN_ERR_ARG, "graph level " + std::to_string(level) + " is not set");
  const auto& G = idx->levels[level];
  const uint64_t chunk = 1u << 20;
  std::vector<uint32_t> buf;
  const uint32_t md = G.max_degree;
  for (uint64_t r0 = 0; r0 < idx->n; r0 += chunk) {
    uint64_t r1 = std::min<uint64_t>(idx->n, r0 + chunk);
    buf.resize((r1 - r0) * G.nbr_stride);
    HIP_TRY(hipMemcpy(buf.data(), G.d_nbrs + r0 * G.nbr_stride, buf.size() * 4,
                      hipMemcpyDeviceToHost),
            "hipMemcpy(graph)");
    for (uint64_t i = r0; i < r1; i++) {
      const uint32_t* src = &buf[(i - r0) * G.nbr_stride];
      uint32_t* dst = graph_host + i * (md + 1);
      uint32_t d = 0;
      while (d < md && src[d] != PANN_SENTINEL) d++;
      dst[0] = d;
      std::memcpy(dst + 1, src, d * 4);
      std::memset(dst + 1 + d, 0, (md - d) * 4);
    }
  }
  return PANN_OK;
}

}  // namespace

pann_status pann_update_neighbors(pann_index_t idx, const uint32_t* pts, uint32_t np,
                                  const uint32_t* lists, const uint32_t* counts, uint32_t R,
                                  void* stream) {
  if (!idx || !pts || !lists || !counts) return fail(PANN_ERR_ARG, "null argument");
  if (np == 0) return PANN_OK;
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  return update_level(idx, 0, pts, np, lists, counts, R, (hipStream_t)stream);
}

pann_status pann_download_graph(pann_index_t idx, uint32_t* graph_host) {
  if (!idx || !graph_host) return fail(PANN_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  return download_level(idx, 0, graph_host);
}

/* ------------------------------------------------------------------------ */
/* HNSW (algorithms/HNSW/HNSW.hpp)                                           */
/* ------------------------------------------------------------------------ */

pann_status pann_hnsw_set_level(pann_index_t idx, uint32_t level, uint32_t max_degree,
                                const uint32_t* graph_host) {
  if (!idx) return fail(PANN_ERR_ARG, "null index");
  if (level == 0) return fail(PANN_ERR_ARG, "level 0 is the graph given to pann_open");
  if (level > 64) return fail(PANN_ERR_ARG, "level > 64");
  if (max_degree == 0 || max_degree > 64)
    return fail(PANN_ERR_UNSUPPORTED, "level max_degree must be in [1, 64]");
  if (graph_host) {
    pann_status cst = check_graph(idx->n, max_degree, graph_host);
    if (cst != PANN_OK) return cst;
  }
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  if (idx->levels.size() <= level) idx->levels.resize(level + 1);
  auto& G = idx->levels[level];
  if (G.d_nbrs) (void)hipFree(G.d_nbrs);
  G = pann_index_s::Level{};
  const uint32_t stride = (uint32_t)round_up(max_degree, 16);
  uint32_t* d = nullptr;
  HIP_TRY(dev_malloc(&d, (size_t)idx->n * stride * 4), "hipMalloc(level graph)");
  pann_status st = upload_graph(idx->n, max_degree, stride, graph_host, d);
  if (st != PANN_OK) {
    (void)hipFree(d);
    return st;
  }
  G.d_nbrs = d;
  G.max_degree = max_degree;
  G.nbr_stride = stride;
  return PANN_OK;
}

static pann_status hnsw_search_layer_once(pann_index_t idx, uint32_t level, const void* queries,
                                          const uint32_t* self_ids, uint32_t nq, uint32_t ef,
                                          uint32_t degree_limit, int64_t limit,
                                          const uint32_t* starts, const uint32_t* nstarts,
                                          uint32_t starts_stride, const pann_search_opts* opts,
                                          uint32_t* out_ids, float* out_dists, uint32_t out_stride,
                                          uint32_t* out_count, uint32_t* out_visited,
                                          uint32_t* out_cmps, void* stream_) {
  if (!idx || !starts || !out_ids || !out_dists || !out_count)
    return fail(PANN_ERR_ARG, "null argument");
  if ((queries == nullptr) == (self_ids == nullptr))
    return fail(PANN_ERR_ARG, "give exactly one of queries and self_ids");
  if (nq == 0) return PANN_OK;
  if (ef == 0 || out_stride == 0) return fail(PANN_ERR_ARG, "ef and out_stride must be >= 1");
  if (starts_stride == 0 && nstarts) return fail(PANN_ERR_ARG, "starts_stride must be >= 1");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  hipStream_t stream = (hipStream_t)stream_;
  // the intended QueryParams(ef, ef, limit, threshold_m(l)) of search_layer
  // (HNSW.hpp:1096); the reference's literal 1.35 lands in `limit`
  pann_qparams qp{(int64_t)std::min(ef, out_stride), (int64_t)ef,
                  limit > 0 ? limit : (int64_t)idx->n, (int64_t)degree_limit, 100.0, 0.125};
  StartSets ss{starts, nstarts, nstarts ? starts_stride : 1u};
  pann::SearchArgs a;
  int store_kind;
  bool sequential;
  uint32_t grid;
  StreamCtx* C = nullptr;
  pann_status st = prepare_search(idx, level, &qp, nullptr, 0, &ss, opts, nq, a, store_kind,
                                  sequential, grid, stream, C);
  if (st != PANN_OK) return st;
  const int32_t store = opts ? opts->store : PANN_STORE_NATIVE;
  if (self_ids) {  // query q is point self_ids[q]; same_as skips it
    a.queries = store == PANN_STORE_Q8 ? idx->d_q8 : idx->d_pts;
    a.q_stride = store == PANN_STORE_Q8 ? idx->q8_stride : idx->pts_stride;
    a.fqueries = idx->d_pts;
    a.fq_stride = idx->pts_stride;
    a.self_ids = self_ids;
  } else {
    const bool f32_queries = idx->dtype == PANN_DT_F32;
    a.queries = (const char*)queries;
    a.q_stride = (uint64_t)idx->dim * (f32_queries ? 4 : 1);
    a.fqueries = (const char*)queries;
    a.fq_stride = (uint64_t)idx->dim * 4;
    st = stage_queries(idx, *C, store, queries, nq, a, stream);
    if (st != PANN_OK) return st;
  }
  a.nq = nq;
  a.out_ids = out_ids;
  a.out_dists = out_dists;
  a.out_stride = out_stride;
  a.out_visited = out_visited;
  a.out_cmps = out_cmps;
  a.out_nfront = out_count;
  a.short_ok = 1u;  // a layer returns whatever frontier it reaches
  grid = std::min<uint32_t>(grid, nq);
  HIP_TRY(pann::launch_beam_search(a, store_kind, idx->metric, sequential, grid, stream),
          "search_layer launch");
  if (opts && opts->async) return PANN_OK;
  return check_status(*C, stream, true);
}

pann_status pann_hnsw_search_layer(pann_index_t idx, uint32_t level, const void* queries,
                                   const uint32_t* self_ids, uint32_t nq, uint32_t ef,
                                   uint32_t degree_limit, int64_t limit, const uint32_t* starts,
                                   const uint32_t* nstarts, uint32_t starts_stride,
                                   const pann_search_opts* opts, uint32_t* out_ids,
                                   float* out_dists, uint32_t out_stride, uint32_t* out_count,
                                   uint32_t* out_visited, uint32_t* out_cmps, void* stream_) {
  return grow_retry(idx, [&] {
    return hnsw_search_layer_once(idx, level, queries, self_ids, nq, ef, degree_limit, limit,
                                  starts, nstarts, starts_stride, opts, out_ids, out_dists,
                                  out_stride, out_count, out_visited, out_cmps, stream_);
  });
}

static pann_status hnsw_search_once(pann_index_t idx, const void* queries, uint32_t nq, uint32_t k,
                                    uint32_t ef, const int32_t* k_pq, const int32_t* ef_pq,
                                    int64_t limit, const uint32_t* entrance, uint32_t nentrance,
                                    uint32_t top_level, const uint32_t* node_ids,
                                    const pann_search_opts* opts, uint32_t* out_ids,
                                    float* out_dists, uint32_t out_stride, uint32_t* out_count,
                                    uint32_t* out_visited, uint32_t* out_cmps, void* stream_) {
  if (!idx || !queries || !entrance || !out_ids || !out_dists || !out_count)
    return fail(PANN_ERR_ARG, "null argument");
  if (nq == 0) return PANN_OK;
  if (ef == 0) return fail(PANN_ERR_ARG, "ef must be >= 1");
  if (out_stride < ef) return fail(PANN_ERR_ARG, "out_stride smaller than ef");
  if (nentrance == 0) return fail(PANN_ERR_ARG, "empty entrance");
  if (top_level >= idx->levels.size()) return fail(PANN_ERR_ARG, "top_level is not set");
  for (uint32_t l = 0; l <= top_level; l++)
    if (!idx->levels[l].d_nbrs)
      return fail(PANN_ERR_ARG, "graph level " + std::to_string(l) + " is not set");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  hipStream_t stream = (hipStream_t)stream_;
  StreamCtx* C = nullptr;
  {
    pann_status cst = stream_ctx(idx, stream, &C);
    if (cst != PANN_OK) return cst;
  }
  if (C->eps_cap < nq) {
    if (C->d_eps || C->d_epsd) HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (C->d_eps) (void)hipFree(C->d_eps);
    if (C->d_epsd) (void)hipFree(C->d_epsd);
    C->d_eps = nullptr;
    C->d_epsd = nullptr;
    C->eps_cap = 0;
    HIP_TRY(dev_malloc(&C->d_eps, (size_t)nq * 4), "hipMalloc(entry points)");
    HIP_TRY(dev_malloc(&C->d_epsd, (size_t)nq * 4), "hipMalloc(entry points)");
    C->eps_cap = nq;
  }
  pann_search_opts o{};
  if (opts) o = *opts;
  const bool f32_queries = idx->dtype == PANN_DT_F32;
  StartSets ss{C->d_eps, nullptr, 1u};
  HIP_TRY(hipEventRecord(C->ev0, stream), "hipEventRecord");
  if (top_level <= pann::PANN_MAX_UP) {
    // one launch: each wave descends its query through levels top .. 1 at
    // beam 1 and then searches layer 0 with ef, so the descents of some
    // queries overlap the layer-0 searches of others
    const auto& G = idx->levels[0];
    pann_qparams qp{(int64_t)ef, (int64_t)ef, limit > 0 ? limit : (int64_t)idx->n,
                    (int64_t)G.max_degree, 100.0, 0.125};
    pann::SearchArgs a;
    int store_kind;
    bool sequential;
    uint32_t grid;
    StreamCtx* Cl = nullptr;
    pann_status st = prepare_search(idx, 0, &qp, entrance, nentrance, nullptr, &o, nq, a,
                                    store_kind, sequential, grid, stream, Cl, top_level + 1,
                                    nullptr, !(ef_pq && !o.query_order));
    if (st != PANN_OK) return st;
    a.queries = (const char*)queries;
    a.q_stride = (uint64_t)idx->dim * (f32_queries ? 4 : 1);
    a.fqueries = (const char*)queries;
    a.fq_stride = (uint64_t)idx->dim * 4;
    st = stage_queries(idx, *C, o.store, queries, nq, a, stream);
    if (st != PANN_OK) return st;
    a.nq = nq;
    a.out_ids = out_ids;
    a.out_dists = out_dists;
    a.out_stride = out_stride;
    a.out_nfront = out_count;
    a.out_visited = out_visited;
    a.out_cmps = out_cmps;  // count_cmps sums every level (:1553, :1097-1098)
    a.accumulate = 0u;
    a.short_ok = 1u;
    if (ef_pq) {  // per-query ef: the whole frontier comes back
      a.k_pq = ef_pq;
      a.beam_pq = ef_pq;
      if (!o.query_order) {
        st = order_by_beam(*C, ef_pq, nq, ef, a, stream);
        if (st != PANN_OK) return st;
      }
    }
    a.nup = top_level;
    for (uint32_t i = 0; i < top_level; i++) {
      const auto& U = idx->levels[top_level - i];
      a.up[i] = pann::UpperLevel{U.d_nbrs, U.nbr_stride, U.max_degree};
    }
    grid = std::min<uint32_t>(grid, nq);
    HIP_TRY(pann::launch_beam_search(a, store_kind, idx->metric, sequential, grid, stream),
            "HNSW search launch");
    HIP_TRY(pann::launch_hnsw_cut(out_ids, out_dists, out_stride, out_count, nq, k, k_pq,
                                  node_ids, stream),
            "HNSW cut launch");
    HIP_TRY(hipEventRecord(C->ev1, stream), "hipEventRecord");
    idx->last_timed = C;
    if (o.async) return PANN_OK;
    return check_status(*C, stream, true);
  }
  // more levels than one launch takes: a launch per level
  // search_layer_to (HNSW.hpp:1545-1566): ef = 1 from the entrance through
  // levels top .. 1, keeping the closest; then layer 0 with ef (:1582-1586)
  for (int64_t l = top_level; l >= 0; l--) {
    const auto& G = idx->levels[(size_t)l];
    const bool last = l == 0;
    const uint32_t beam = last ? ef : 1u;
    pann_qparams qp{(int64_t)beam, (int64_t)beam,
                    last && limit > 0 ? limit : (int64_t)idx->n, (int64_t)G.max_degree,
                    100.0, 0.125};
    pann::SearchArgs a;
    int store_kind;
    bool sequential;
    uint32_t grid;
    pann_search_opts lo = o;
    lo.rerank = last ? o.rerank : 0;
    const bool first = l == (int64_t)top_level;
    StreamCtx* Cl = nullptr;
    pann_status st = prepare_search(idx, (uint32_t)l, &qp, entrance, nentrance,
                                    first ? nullptr : &ss, &lo, nq, a, store_kind, sequential,
                                    grid, stream, Cl, 1, nullptr,
                                    !(last && ef_pq && !o.query_order));
    if (st != PANN_OK) return st;
    a.queries = (const char*)queries;
    a.q_stride = (uint64_t)idx->dim * (f32_queries ? 4 : 1);
    a.fqueries = (const char*)queries;
    a.fq_stride = (uint64_t)idx->dim * 4;
    if (first) {
      st = stage_queries(idx, *C, o.store, queries, nq, a, stream);
      if (st != PANN_OK) return st;
    } else if (o.store == PANN_STORE_Q8) {
      a.queries = (const char*)C->d_qcodes;
      a.q_stride = idx->dim;
    }
    a.nq = nq;
    a.out_ids = last ? out_ids : C->d_eps;  // in place: each wave reads its start first
    a.out_dists = last ? out_dists : C->d_epsd;
    a.out_stride = last ? out_stride : 1u;
    a.out_nfront = last ? out_count : nullptr;
    a.out_visited = out_visited;
    a.out_cmps = out_cmps;  // count_cmps sums every level (:1553, :1097-1098)
    a.accumulate = first ? 0u : 1u;
    a.short_ok = 1u;
    if (last && ef_pq) {  // per-query ef: the whole frontier comes back
      a.k_pq = ef_pq;
      a.beam_pq = ef_pq;
      if (!o.query_order) {
        st = order_by_beam(*C, ef_pq, nq, ef, a, stream);
        if (st != PANN_OK) return st;
      }
    }
    grid = std::min<uint32_t>(grid, nq);
    HIP_TRY(pann::launch_beam_search(a, store_kind, idx->metric, sequential, grid, stream),
            "HNSW layer launch");
  }
  HIP_TRY(pann::launch_hnsw_cut(out_ids, out_dists, out_stride, out_count, nq, k, k_pq,
                                node_ids, stream),
          "HNSW cut launch");
  HIP_TRY(hipEventRecord(C->ev1, stream), "hipEventRecord");
  idx->last_timed = C;
  if (o.async) return PANN_OK;
  return check_status(*C, stream, true);
}

pann_status pann_hnsw_search(pann_index_t idx, const void* queries, uint32_t nq, uint32_t k,
                             uint32_t ef, const int32_t* k_pq, const int32_t* ef_pq,
                             int64_t limit, const uint32_t* entrance,
                             uint32_t nentrance, uint32_t top_level, const uint32_t* node_ids,
                             const pann_search_opts* opts, uint32_t* out_ids, float* out_dists,
                             uint32_t out_stride, uint32_t* out_count, uint32_t* out_visited,
                             uint32_t* out_cmps, void* stream_) {
  return grow_retry(idx, [&] {
    return hnsw_search_once(idx, queries, nq, k, ef, k_pq, ef_pq, limit, entrance, nentrance,
                            top_level, node_ids, opts, out_ids, out_dists, out_stride, out_count,
                            out_visited, out_cmps, stream_);
  });
}

pann_status pann_hnsw_select_neighbors(pann_index_t idx, uint32_t level, const uint32_t* pts,
                                       uint32_t np, const uint32_t* cand_ids,
                                       const float* cand_dists, const uint32_t* cand_count,
                                       uint32_t cand_stride, float alpha, uint32_t M,
                                       uint32_t* out_ids, uint32_t* out_count, void* stream) {
  if (!idx || !pts || !cand_ids || !cand_dists || !cand_count || !out_ids || !out_count)
    return fail(PANN_ERR_ARG, "null argument");
  if (np == 0) return PANN_OK;
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  return run_prune(idx, level, pts, np, cand_ids, cand_dists, cand_count, cand_stride, nullptr,
                   cand_stride, (double)alpha, M, 4, out_ids, out_count, (hipStream_t)stream);
}

pann_status pann_hnsw_add_reverse(pann_index_t idx, uint32_t level, const uint32_t* pts,
                                  uint32_t np, const uint32_t* offsets, const uint32_t* cand_ids,
                                  uint32_t max_count, uint32_t M, uint32_t* out_ids,
                                  uint32_t* out_count, void* stream) {
  if (!idx || !pts || !offsets || !cand_ids || !out_ids || !out_count)
    return fail(PANN_ERR_ARG, "null argument");
  if (np == 0) return PANN_OK;
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  return run_prune(idx, level, pts, np, cand_ids, nullptr, nullptr, 0, offsets, max_count, 1.0,
                   M, 5, out_ids, out_count, (hipStream_t)stream);
}

pann_status pann_hnsw_set_neighbors(pann_index_t idx, uint32_t level, const uint32_t* pts,
                                    uint32_t np, const uint32_t* lists, const uint32_t* counts,
                                    uint32_t R, void* stream) {
  if (!idx || !pts || !lists || !counts) return fail(PANN_ERR_ARG, "null argument");
  if (np == 0) return PANN_OK;
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  return update_level(idx, level, pts, np, lists, counts, R, (hipStream_t)stream);
}

pann_status pann_hnsw_download_level(pann_index_t idx, uint32_t level, uint32_t* graph_host) {
  if (!idx || !graph_host) return fail(PANN_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  return download_level(idx, level, graph_host);
}

// The brute-force candidates of nq queries (rows qs floats apart) among the
// n rows of X (point dtype pdt, rows xs elements apart, xnorm = |x|^2 for
// L2): the kk best per query, sorted.  Synchronises `stream`.
static pann_status run_bruteforce(const pann_index_s* idx, const void* X, int pdt, uint64_t xs,
                                  const float* xnorm, uint32_t n, int metric, const float* Q,
                                  uint64_t qs, uint32_t nq, uint32_t kk, uint32_t* out_ids,
                                  float* out_dists, hipStream_t stream) {
  pann::BFArgs a;
  std::memset(&a, 0, sizeof(a));
  a.X = X;
  a.xs = xs;
  a.xnorm = metric == PANN_METRIC_L2 ? xnorm : nullptr;
  a.n = n;
  a.dim = idx->dim;
  a.dim_pad = pann::bf_dim_pad(idx->dim, pdt);  // inside the 128-byte padded rows
  a.qs = qs;
  a.metric = metric == PANN_METRIC_L2 ? 0 : 1;
  a.kk = kk;
  a.cap = 1;
  while (a.cap < kk + 32) a.cap <<= 1;
  // point ranges: a multiple of 8 (the XCD count), so that blockIdx
  // (query tile * nranges + range) puts every wave of a range on one XCD and
  // its L2 serves the range's rows to all of them; more ranges for few
  // queries, each range >= 4096 points, and nranges * kk within the merge
  const uint32_t tiles = (nq + 63) / 64;
  uint32_t nr = std::max<uint32_t>(8, (uint32_t)(idx->num_cus * 8 + tiles - 1) / tiles);
  nr = std::min<uint32_t>((nr + 7) & ~7u, 64);
  nr = std::min<uint32_t>(nr, std::max<uint32_t>(1, n / 4096));
  nr = std::min<uint32_t>(nr, 4096 / kk);
  a.nranges = std::max<uint32_t>(1, nr);
  a.range_len = (n + a.nranges - 1) / a.nranges;
  // queries per launch: scratch (cap keys per query and range) within 1 GiB
  const uint64_t per_q = (uint64_t)a.nranges * (a.cap * 8 + 4);
  const uint32_t chunk = (uint32_t)std::max<uint64_t>(
      64, std::min<uint64_t>(nq, ((1ull << 30) / per_q) & ~63ull));
  uint64_t* buf = nullptr;
  uint32_t* counts = nullptr;
  const uint32_t cq = std::min(nq, chunk);
  hipError_t e = dev_malloc(&buf, (size_t)cq * a.nranges * a.cap * 8);
  if (e == hipSuccess) e = dev_malloc(&counts, (size_t)cq * a.nranges * 4);
  if (e != hipSuccess) {
    if (buf) (void)hipFree(buf);
    return hip_fail(e, "hipMalloc(brute-force scratch)");
  }
  a.buf = buf;
  a.counts = counts;
  for (uint32_t q0 = 0; q0 < nq && e == hipSuccess; q0 += chunk) {
    a.nq = std::min(chunk, nq - q0);
    a.Q = Q + (uint64_t)q0 * qs;
    e = pann::launch_bf(a, pdt, out_ids + (uint64_t)q0 * kk, out_dists + (uint64_t)q0 * kk, stream);
  }
  hipError_t e2 = hipStreamSynchronize(stream);
  (void)hipFree(buf);
  (void)hipFree(counts);
  if (e != hipSuccess) return hip_fail(e, "brute-force launch");
  if (e2 != hipSuccess) return hip_fail(e2, "hipStreamSynchronize");
  return PANN_OK;
}

pann_status pann_bruteforce_candidates(pann_index_t idx, const float* queries, uint32_t nq,
                                      uint32_t kk, uint32_t* out_ids, float* out_dists,
                                      void* stream_) {
  if (!idx || !queries || !out_ids || !out_dists) return fail(PANN_ERR_ARG, "null argument");
  if (nq == 0) return PANN_OK;
  if (kk == 0 || kk > 224) return fail(PANN_ERR_ARG, "kk must be in [1, 224]");
  if (idx->dim > 256)
    return fail(PANN_ERR_UNSUPPORTED, "brute force: dim > 256");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  hipStream_t stream = (hipStream_t)stream_;
  const int pdt = idx->dtype == PANN_DT_F32 ? 0 : (idx->dtype == PANN_DT_U8 ? 1 : 2);
  const uint64_t xs = idx->pts_stride / elem_bytes(idx->dtype);
  if (idx->metric == PANN_METRIC_L2 && !idx->d_xnorm) {
    HIP_TRY(dev_malloc(&idx->d_xnorm, (size_t)idx->n * 4), "hipMalloc(norms)");
    HIP_TRY(pann::launch_row_sqnorm(idx->d_pts, pdt, xs, idx->n, idx->dim, idx->d_xnorm, stream),
            "norms launch");
  }
  return run_bruteforce(idx, idx->d_pts, pdt, xs, idx->d_xnorm, idx->n, idx->metric, queries,
                        idx->dim, nq, kk, out_ids, out_dists, stream);
}

pann_status pann_build_fp8_residual(pann_index_t idx, uint32_t ncent) {
  if (!idx) return fail(PANN_ERR_ARG, "null index");
  if (idx->dtype != PANN_DT_F32)
    return fail(PANN_ERR_ARG, "reduced-precision stores need float points");
  if (idx->metric != PANN_METRIC_MIPS)
    return fail(PANN_ERR_UNSUPPORTED, "the FP8 residual store is for inner-product indexes");
  if (idx->dim > 256) return fail(PANN_ERR_UNSUPPORTED, "FP8 residual store: dim > 256");
  if (idx->n < 4) return fail(PANN_ERR_UNSUPPORTED, "FP8 residual store: fewer than 4 points");
  if (ncent == 0) ncent = std::max<uint32_t>(4, std::min<uint32_t>(2048, (idx->n / 64) & ~3u));
  if (ncent < 4 || ncent % 4 != 0 || ncent > 4096 || ncent > idx->n)
    return fail(PANN_ERR_ARG, "ncent must be a multiple of 4 in [4, min(4096, n)]");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  void* old[] = {idx->d_fp8r, idx->d_rscales, idx->d_rabsmax, idx->d_cent,
                 idx->d_assign};
  for (void* p : old)
    if (p) (void)hipFree(p);
  idx->d_fp8r = nullptr;
  idx->d_rscales = nullptr;
  idx->d_rabsmax = nullptr;
  idx->d_cent = nullptr;
  idx->d_assign = nullptr;
  idx->ncent = 0;
  const uint64_t ps = idx->pts_stride, fs = ps / 4;
  const uint32_t n = idx->n, dim = idx->dim;
  idx->fp8r_stride = round_up((uint64_t)dim + 4, 128);
  HIP_TRY(dev_malloc(&idx->d_fp8r, (size_t)n * idx->fp8r_stride), "hipMalloc(fp8 residual)");
  HIP_TRY(hipMemset(idx->d_fp8r, 0, (size_t)n * idx->fp8r_stride), "hipMemset(fp8 residual)");
  HIP_TRY(dev_malloc(&idx->d_rscales, dim * 4), "hipMalloc(scales)");
  HIP_TRY(dev_malloc(&idx->d_rabsmax, dim * 4), "hipMalloc(absmax)");
  HIP_TRY(dev_malloc(&idx->d_cent, (size_t)ncent * ps), "hipMalloc(centroids)");
  HIP_TRY(dev_malloc(&idx->d_assign, (size_t)n * 4), "hipMalloc(assignment)");
  // centroids: every (n / ncent)-th point
  const uint64_t step = n / ncent;
  HIP_TRY(hipMemcpy2D(idx->d_cent, ps, idx->d_pts, ps * step, ps, ncent, hipMemcpyDeviceToDevice),
          "hipMemcpy2D(centroids)");
  // each point's nearest centroid in L2, on the brute-force kernel
  float* cnorm = nullptr;
  float* cd = nullptr;
  hipError_t e = dev_malloc(&cnorm, (size_t)ncent * 4);
  if (e == hipSuccess) e = dev_malloc(&cd, (size_t)n * 4);
  if (e == hipSuccess) e = pann::launch_row_sqnorm(idx->d_cent, 0, fs, ncent, dim, cnorm, nullptr);
  pann_status st = e == hipSuccess ? PANN_OK : hip_fail(e, "centroid norms");
  if (st == PANN_OK)
    st = run_bruteforce(idx, idx->d_cent, 0, fs, cnorm, ncent, PANN_METRIC_L2,
                        (const float*)idx->d_pts, fs, n, 1, idx->d_assign, cd, nullptr);
  if (cnorm) (void)hipFree(cnorm);
  if (cd) (void)hipFree(cd);
  if (st != PANN_OK) return st;
  HIP_TRY(pann::launch_fp8_store((const float*)idx->d_pts, fs, n, dim, idx->d_rabsmax,
                                 idx->d_rscales, (uint8_t*)idx->d_fp8r, idx->fp8r_stride, nullptr,
                                 idx->d_cent, fs, idx->d_assign),
          "fp8 residual store kernel");
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  idx->ncent = ncent;
  return PANN_OK;
}

pann_status pann_get_fp8_residual(pann_index_t idx, uint32_t* ncent, float* centroids,
                                  uint32_t* assign, float* scales) {
  if (!idx || !ncent) return fail(PANN_ERR_ARG, "null argument");
  if (!idx->d_fp8r || idx->ncent == 0) return fail(PANN_ERR_ARG, "FP8 residual store not built");
  std::lock_guard<std::mutex> lk(idx->mu);
  HIP_TRY(hipSetDevice(idx->device), "hipSetDevice");
  *ncent = idx->ncent;
  if (centroids)
    HIP_TRY(hipMemcpy2D(centroids, (size_t)idx->dim * 4, idx->d_cent, idx->pts_stride,
                        (size_t)idx->dim * 4, idx->ncent, hipMemcpyDeviceToHost),
            "hipMemcpy2D(centroids)");
  if (assign)
    HIP_TRY(hipMemcpy(assign, idx->d_assign, (size_t)idx->n * 4, hipMemcpyDeviceToHost),
            "hipMemcpy(assignment)");
  if (scales)
    HIP_TRY(hipMemcpy(scales, idx->d_rscales, (size_t)idx->dim * 4, hipMemcpyDeviceToHost),
            "hipMemcpy(scales)");
  return PANN_OK;
}

pann_status pann_close(pann_index_t idx) {
  if (!idx) return PANN_OK;
  (void)hipSetDevice(idx->device);
  (void)hipDeviceSynchronize();
  void* ptrs[] = {idx->d_pts,    idx->d_bf16,    idx->d_fp8,   idx->d_scales, idx->d_absmax,
                  idx->d_nbrs,   idx->d_q8,      idx->d_xnorm, idx->d_dummy,  idx->d_pkeys,
                  idx->d_pflags, idx->d_fp8r,    idx->d_rscales, idx->d_rabsmax, idx->d_cent,
                  idx->d_assign, idx->d_evals, idx->d_fp8_bound, idx->d_q8_bound,
                  idx->d_bf16_bound};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (size_t l = 1; l < idx->levels.size(); l++)
    if (idx->levels[l].d_nbrs) (void)hipFree(idx->levels[l].d_nbrs);
  for (auto& kv : idx->streams) {
    StreamCtx& C = kv.second;
    void* cp[] = {C.d_counter, C.d_status, C.d_starts, C.d_ghash, C.d_vis, C.d_eps, C.d_epsd,
                  C.d_qcodes, C.d_qtable, C.d_stage, C.d_cmps, C.d_order};
    for (void* p : cp)
      if (p) (void)hipFree(p);
    if (C.ev0) (void)hipEventDestroy(C.ev0);
    if (C.ev1) (void)hipEventDestroy(C.ev1);
  }
  delete idx;
  return PANN_OK;
}

}  // extern "C"
