// beam_search.hip -- host-side launch of the beam-search kernels
// (beam_search_kernel.h), whose instantiations are spread over one
// translation unit per store (bs_*.hip).
#include "beam_search_kernel.h"

#include <algorithm>

namespace pann {


// Longest-first order of a batch with per-query beams.  The persistent grid
// takes queries in index order, so a long search that starts late sets the
// batch's end; a counting sort by descending beam starts the long ones first
// and leaves the short ones for the tail.  Bin b holds beam max_beam - b.
__device__ __forceinline__ uint32_t order_bin(int32_t beam, uint32_t max_beam) {
  const uint32_t b = beam < 0 ? 0u : ((uint32_t)beam > max_beam ? max_beam : (uint32_t)beam);
  return max_beam - b;
}

// Adds `active` lanes to their bins, one atomic per distinct bin of the wave
// (a batch often has a few beam values, whose bins every lane would hit);
// returns each lane's position: its bin's old count plus its rank among the
// wave's lanes with that bin.  Called by all 64 lanes.
__device__ __forceinline__ uint32_t wave_bin_add(uint32_t* bins, uint32_t bin, bool active) {
  uint64_t rem = ballot(active);
  uint32_t pos = 0;
  while (rem) {
    const int leader = __ffsll((long long)rem) - 1;
    const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)bin, leader);
    const uint64_t m = ballot(active && bin == b) & rem;
    uint32_t base = 0;
    if ((int)threadIdx.x % PANN_WAVE == leader) base = atomicAdd(&bins[b], (uint32_t)__popcll(m));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    if ((m >> (threadIdx.x % PANN_WAVE)) & 1u) pos = base + mbcnt(m);
    rem &= ~m;
  }
  return pos;
}

// bins[0 .. nbins) to exclusive prefix sums, in one workgroup of 1024
// threads: each thread sums a contiguous run of bins, the first wave scans the
// 1024 run sums (16 per lane, then across its lanes), and each thread writes
// its run's prefixes
__device__ __forceinline__ void order_scan(uint32_t* bins, uint32_t nbins, uint32_t* part) {
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nbins + 1023u) / 1024u;
  const uint32_t lo = t * per < nbins ? t * per : nbins;
  const uint32_t hi = lo + per < nbins ? lo + per : nbins;
  uint32_t s = 0;
  for (uint32_t i = lo; i < hi; i++) s += bins[i];
  part[t] = s;
  __syncthreads();
  if (t < PANN_WAVE) {
    uint32_t v[16];
    uint32_t tot = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      v[i] = part[t * 16u + i];
      tot += v[i];
    }
    uint32_t inc = tot;  // inclusive scan of the lanes' totals
#pragma unroll
    for (int o = 1; o < PANN_WAVE; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o);
      if ((int)t >= o) inc += y;
    }
    uint32_t run = inc - tot;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      part[t * 16u + i] = run;
      run += v[i];
    }
  }
  __syncthreads();
  uint32_t base = part[t];
  for (uint32_t i = lo; i < hi; i++) {
    const uint32_t c = bins[i];
    bins[i] = base;
    base += c;
  }
  __syncthreads();
}

// The whole sort in one workgroup of 1024 threads, the bins in LDS.  Each
// thread loads its queries' beams 16 at a time before using any of them, so a
// batch of up to 16k queries pays one load latency per phase.
constexpr uint32_t kOrderLdsBins = 8192;
constexpr int kOrderPerThread = 16;
__global__ __launch_bounds__(1024) void k_order_lds(const int32_t* beam_pq, uint32_t nq,
                                                    uint32_t max_beam, uint32_t* order,
                                                    uint32_t* status, uint32_t* counter) {
  __shared__ uint32_t bins[kOrderLdsBins];
  __shared__ uint32_t part[1024];
  if (threadIdx.x == 0) {  // the search launch that follows starts from zero
    *status = 0u;
    *counter = 0u;
  }
  const uint32_t nbins = max_beam + 1u;
  for (uint32_t i = threadIdx.x; i < nbins; i += 1024u) bins[i] = 0u;
  __syncthreads();
  // every wave runs the same rounds (wave_bin_add is wave-wide)
  for (uint32_t c0 = 0; c0 < nq; c0 += 1024u * kOrderPerThread) {
    uint32_t bin[kOrderPerThread];
#pragma unroll
    for (int t = 0; t < kOrderPerThread; t++) {
      const uint32_t i = c0 + (uint32_t)t * 1024u + threadIdx.x;
      bin[t] = i < nq ? order_bin(beam_pq[i], max_beam) : 0u;
    }
#pragma unroll
    for (int t = 0; t < kOrderPerThread; t++) {
      const uint32_t i = c0 + (uint32_t)t * 1024u + threadIdx.x;
      wave_bin_add(bins, bin[t], i < nq);
    }
  }
  __syncthreads();
  order_scan(bins, nbins, part);
  for (uint32_t c0 = 0; c0 < nq; c0 += 1024u * kOrderPerThread) {
    uint32_t bin[kOrderPerThread];
#pragma unroll
    for (int t = 0; t < kOrderPerThread; t++) {
      const uint32_t i = c0 + (uint32_t)t * 1024u + threadIdx.x;
      bin[t] = i < nq ? order_bin(beam_pq[i], max_beam) : 0u;
    }
#pragma unroll
    for (int t = 0; t < kOrderPerThread; t++) {
      const uint32_t i = c0 + (uint32_t)t * 1024u + threadIdx.x;
      const bool act = i < nq;
      const uint32_t pos = wave_bin_add(bins, bin[t], act);
      if (act) order[pos] = i;
    }
  }
}

hipError_t launch_order_by_beam(const int32_t* beam_pq, uint32_t nq, uint32_t max_beam,
                                uint32_t* order, uint32_t* status, uint32_t* counter,
                                hipStream_t s) {
  // (a beam past the bins would not fit a wave's LDS frontier either)
  if (max_beam + 1u > kOrderLdsBins || !status || !counter) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_order_lds, dim3(1), dim3(1024), 0, s, beam_pq, nq, max_beam, order,
                     status, counter);
  return hipGetLastError();
}

uint32_t search_lds_bytes(const SearchArgs& a) {
  return wave_lds_bytes(a.fcap, a.ccap, a.dim, a.hslots_log2, a.rc_log2, a.fb_log2, a.ncent,
                        a.q_own, a.wpq);
}

static int nch_for(int row_bytes) {
  int c = (row_bytes + 127) / 128;
  if (c <= 1) return 1;
  if (c <= 2) return 2;
  if (c <= 3) return 3;
  if (c <= 4) return 4;
  if (c <= 8) return 8;
  return -1;
}

hipError_t launch_beam_search(const SearchArgs& a, int store, int metric, bool sequential,
                              uint32_t grid, hipStream_t s) {
  int nch = nch_for((int)a.row_bytes);
  if (sequential) {
    if (store != ST_F32) return hipErrorInvalidValue;
    if (nch < 0) nch = 8;  // rows past 8 lines: 8 lines at a time, query in LDS
    return launch_bs_f32_inorder(a, metric, nch, grid, s);
  }
  switch (store) {
    case ST_F32: return launch_bs_f32(a, metric, nch, grid, s);
    case ST_U8: return launch_bs_u8(a, metric, nch, grid, s);
    case ST_I8: return launch_bs_i8(a, metric, nch, grid, s);
    case ST_BF16: return launch_bs_bf16(a, metric, nch, grid, s);
    case ST_FP8: return launch_bs_fp8(a, metric, nch, grid, s);
    case ST_FP8R:  // rows of dim + 4 <= 384 bytes, inner product only
      if (metric != M_IP || !a.qtable || a.ncent == 0) return hipErrorInvalidValue;
      return launch_bs_fp8r(a, metric, nch, grid, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pann
