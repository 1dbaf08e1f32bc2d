// Work-queue construction for the batched kernel (device side).
//
//   ck_reset_kernel    clear counters and bin histogram (a fill kernel, so the
//                      sequence can be captured in a hipGraph)
//   ck_prep_kernel     per pair: empty / length-ratio pairs get their final
//                      zero outputs; pairs the batched kernel cannot take
//                      (rows > 8191, band wider than the LDS row buffer) go
//                      to the int32 list; the rest get a cost bin
//                      (stripes x (W + 30), the kernel's step count)
//   ck_scan_kernel     exclusive scan of the histogram, bins in descending
//                      cost order (longest pairs are handed out first)
//   ck_scatter_kernel  counting-sort scatter of pair indices into the queue
#include "ck_internal.h"

namespace ck {

__global__ void ck_reset_kernel(Counters* ctr, int32_t* hist, int32_t nbins) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < nbins) hist[x] = 0;
  if (x < (int)(sizeof(Counters) / sizeof(int32_t)))
    reinterpret_cast<int32_t*>(ctr)[x] = 0;
}

// Cost of a pair in the batched kernel with S lanes per pair: stripes of S
// rows, each W + 2(S-1) steps; cost levels of 2^shift steps, each split in
// four by lb - la (in steps of 16 residues, the last open-ended).  Pairs
// of one wavefront then have similar band windows stripe by stripe (the
// window is cut where the band leaves the matrix, which depends on lb - la),
// so fewer lanes idle while the longest window of the group finishes.
__device__ __forceinline__ int32_t cost_bin(int32_t la, int32_t lb, int32_t W, int32_t S,
                                            int32_t shift) {
  const int32_t stripes = (la + S - 1) / S;
  const int64_t steps = (int64_t)stripes * (W + 2 * (S - 1));
  int64_t b = steps >> shift;
  if (b > NBINS / 4 - 1) b = NBINS / 4 - 1;
  const int32_t sub = min(3, (lb - la) >> 4);
  return (NBINS - 1) - ((int32_t)b * 4 + sub);   // bin 0 = most expensive
}

// mode 0: route (batched queue + int32 re-run list); mode 1: every pair to
// the queue (the plain form consumes the queue in cost order).
__global__ __launch_bounds__(256) void ck_prep_kernel(
    const int32_t* __restrict__ pairs, int64_t M, const int32_t* __restrict__ lengths,
    const int32_t* __restrict__ hints, int32_t band, float threshold, uint32_t flags,
    int32_t w_cap, int32_t mode, int32_t seg_lanes, int32_t bin_shift,
    int32_t* __restrict__ bin_of, int32_t* __restrict__ hist,
    Counters* __restrict__ ctr, int32_t* __restrict__ plain_list,
    float* __restrict__ out_sims, int32_t* __restrict__ out_scores,
    int32_t* __restrict__ out_matches, uint8_t* __restrict__ out_mask) {
  __shared__ int32_t lhist[NBINS];
  for (int x = threadIdx.x; x < NBINS; x += blockDim.x) lhist[x] = 0;
  __syncthreads();
  int32_t skipped = 0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < M;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t lx = lengths[pairs[2 * p]], ly = lengths[pairs[2 * p + 1]];
    const bool lr_fail = (flags & 0x1u) && length_ratio_fails(lx, ly, threshold);
    if (lx == 0 || ly == 0 || lr_fail) {
      out_sims[p] = 0.0f;
      out_scores[p] = 0;
      if (out_matches) out_matches[p] = 0;
      if (out_mask) out_mask[p] = 0;
      bin_of[p] = -1;
      skipped += lr_fail ? 1 : 0;
      continue;
    }
    const int32_t la = lx < ly ? lx : ly;
    const int32_t lb = lx < ly ? ly : lx;
    const int32_t bw = (band <= 0 || lb <= 50) ? lb : band;
    const int64_t W64 = 2 * (int64_t)bw + 1;
    if (mode == 0 && (la > MAX_FAST_ROWS || W64 > w_cap)) {
      bin_of[p] = -1;
      const int32_t slot = atomicAdd(&ctr->plain_len, 1);
      atomicAdd(&ctr->prep_plain, 1);
      plain_list[slot] = (int32_t)p;
      continue;
    }
    const int32_t bin =
        cost_bin(la, lb, (int32_t)(W64 > (1 << 20) ? (1 << 20) : W64), seg_lanes, bin_shift);
    bin_of[p] = bin;
    atomicAdd(&lhist[bin], 1);
  }
  __syncthreads();
  for (int x = threadIdx.x; x < NBINS; x += blockDim.x)
    if (lhist[x]) atomicAdd(&hist[x], lhist[x]);
  if (skipped) atomicAdd(&ctr->skipped, skipped);
}

// One block of 1024 threads; NBINS = 4096 -> 4 bins per thread.
__global__ __launch_bounds__(1024) void ck_scan_kernel(int32_t* __restrict__ hist,
                                                       Counters* __restrict__ ctr) {
  __shared__ int32_t part[1024];
  constexpr int PER = NBINS / 1024;
  const int t = threadIdx.x;
  int32_t v[PER];
  int32_t sum = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) { v[q] = hist[t * PER + q]; sum += v[q]; }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int32_t add = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int32_t run = part[t] - sum;
#pragma unroll
  for (int q = 0; q < PER; ++q) { hist[t * PER + q] = run; run += v[q]; }
  if (t == 1023) ctr->queue_len = part[1023];
}

// Pairs are scattered in chunks of SCATTER_CHUNK: a block ranks its chunk's
// pairs per bin in LDS, reserves each bin's range with one global atomic,
// then writes.  (One global atomic per pair serialises on the few bins that
// hold most pairs.)  Order inside a bin is unspecified.
constexpr int SCATTER_PER_THREAD = 8;
constexpr int SCATTER_CHUNK = 256 * SCATTER_PER_THREAD;

__global__ __launch_bounds__(256) void ck_scatter_kernel(
    const int32_t* __restrict__ bin_of, int64_t M, int32_t* __restrict__ cursor,
    int32_t* __restrict__ queue) {
  __shared__ int32_t lh[NBINS];
  for (int64_t c0 = (int64_t)blockIdx.x * SCATTER_CHUNK; c0 < M;
       c0 += (int64_t)gridDim.x * SCATTER_CHUNK) {
    for (int x = threadIdx.x; x < NBINS; x += blockDim.x) lh[x] = 0;
    __syncthreads();
    int32_t bin[SCATTER_PER_THREAD], rank[SCATTER_PER_THREAD];
#pragma unroll
    for (int u = 0; u < SCATTER_PER_THREAD; ++u) {
      const int64_t p = c0 + u * 256 + threadIdx.x;
      bin[u] = p < M ? bin_of[p] : -1;
      rank[u] = bin[u] >= 0 ? atomicAdd(&lh[bin[u]], 1) : 0;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < NBINS; x += blockDim.x) {
      const int32_t n = lh[x];
      if (n) lh[x] = atomicAdd(&cursor[x], n);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < SCATTER_PER_THREAD; ++u) {
      const int64_t p = c0 + u * 256 + threadIdx.x;
      if (bin[u] >= 0) queue[lh[bin[u]] + rank[u]] = (int32_t)p;
    }
    __syncthreads();
  }
}

}  // namespace ck

namespace ck {

// Search stage 2 input: pairs[q*mc + c] = (q_index_base + q, targets[q*mc + c])
// for c < counts[q], else the empty-sequence pair, which the SW entry point
// answers with zeros without aligning.
__global__ __launch_bounds__(256) void ck_search_pairs_kernel(
    const int32_t* __restrict__ targets, const int32_t* __restrict__ counts, int32_t nq,
    int32_t mc, int32_t q_index_base, int32_t empty_index, int32_t* __restrict__ pairs) {
  const int64_t n = (int64_t)nq * mc;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int32_t q = (int32_t)(x / mc), c = (int32_t)(x % mc);
    const bool valid = c < counts[q];
    pairs[2 * x] = valid ? q_index_base + q : empty_index;
    pairs[2 * x + 1] = valid ? targets[x] : empty_index;
  }
}

}  // namespace ck
