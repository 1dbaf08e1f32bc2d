// Top-k hits per query on the device (clustkit/search.py:727-800 with the
// rank score of clustkit/kmer_index.py:2318-2322).
//
// Input per query q: candidate slots 0 .. counts[q]-1 of targets[q*mc ..],
// with SW identity, score and pass mask per slot.  A slot enters when it is
// below the count and its mask is set.  Rank (rank_mode): 0 identity, 2 the
// score as float32, 1 rank = fl32(fl32(0.267f * fl32(score)) - logtab[qlen*tlen])
// where logtab[p] is the float32 natural log of p as NumPy computes it for a
// float32 array (the host fills the table with NumPy itself; every product
// below 2^24 is an exact float32 integer, so the table index is exact).  The
// multiply and the subtraction are rounded separately (floating-point
// contraction off), as NumPy does them.
//
// Order: rank descending, then slot ascending -- NumPy's stable lexsort by
// (query, -rank) over the passing pairs in (query, slot) order; the first
// top_k of each query are written, *out_n[q] of them.
//
// One 256-thread workgroup per query: 64-bit keys (order-preserving rank
// bits, inverted for descending, above the slot) bitonic-sorted in LDS.
#include "ck_internal.h"

namespace ck {

namespace {

constexpr int TK_BLOCK = 256;
constexpr int TK_MAX = 2048;   // candidates per query held in LDS

__device__ __forceinline__ uint32_t float_order(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);   // ascending float order
}

__global__ __launch_bounds__(TK_BLOCK) void topk_kernel(
    const int32_t* __restrict__ targets, const int32_t* __restrict__ counts,
    const float* __restrict__ sims, const int32_t* __restrict__ scores,
    const uint8_t* __restrict__ mask, const int32_t* __restrict__ qlens,
    const int32_t* __restrict__ tlens, const float* __restrict__ logtab, int64_t logtab_len,
    int32_t nq, int32_t mc, int32_t top_k, int32_t rank_mode, int32_t* __restrict__ out_target,
    float* __restrict__ out_identity, float* __restrict__ out_rank,
    int32_t* __restrict__ out_n, int32_t* __restrict__ status) {
  __shared__ uint64_t key[TK_MAX];
  __shared__ float rk[TK_MAX];
  __shared__ int32_t nvalid;
  int P = 64;
  while (P < mc) P <<= 1;
  for (int32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    if (threadIdx.x == 0) nvalid = 0;
    __syncthreads();
    const int32_t cnt = min(counts[q], mc);
    const int64_t base = (int64_t)q * mc;
    for (int x = threadIdx.x; x < P; x += TK_BLOCK) {
      uint64_t k = ~0ull;
      if (x < cnt && mask[base + x]) {
        float r = sims[base + x];
        if (rank_mode == 2) {
          r = (float)scores[base + x];   // the raw score, as NumPy's astype(float32)
        } else if (rank_mode == 1) {
          const int64_t prod = (int64_t)qlens[q] * (int64_t)tlens[targets[base + x]];
          if (prod < 1 || prod >= logtab_len) {
            atomicOr(status, 1);       // the host sized the table too small
            r = 0.0f;
          } else {
            // Two roundings, as NumPy does them: no fused multiply-add.
#pragma clang fp contract(off)
            const float m = 0.267f * (float)scores[base + x];
            r = m - logtab[prod];
          }
        }
        rk[x] = r;
        // Descending rank, then ascending slot.
        k = ((uint64_t)(~float_order(r)) << 32) | (uint32_t)x;
        atomicAdd(&nvalid, 1);
      }
      key[x] = k;
    }
    __syncthreads();
    for (int kk = 2; kk <= P; kk <<= 1) {
      for (int j = kk >> 1; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < (P >> 1); t += TK_BLOCK) {
          const int i = 2 * t - (t & (j - 1));
          const int l = i + j;
          const uint64_t a = key[i], b = key[l];
          const bool up = (i & kk) == 0;
          if ((a > b) == up) {
            key[i] = b;
            key[l] = a;
          }
        }
        __syncthreads();
      }
    }
    const int32_t nout = min(nvalid, top_k);
    for (int x = threadIdx.x; x < nout; x += TK_BLOCK) {
      const int32_t slot = (int32_t)(uint32_t)key[x];
      const int64_t o = (int64_t)q * top_k + x;
      out_target[o] = targets[base + slot];
      out_identity[o] = sims[base + slot];
      out_rank[o] = rk[slot];
    }
    if (threadIdx.x == 0) out_n[q] = nout;
    __syncthreads();
  }
}

}  // namespace

int32_t topk_max_candidates() { return TK_MAX; }

hipError_t search_topk(const int32_t* targets, const int32_t* counts, const float* sims,
                       const int32_t* scores, const uint8_t* mask, const int32_t* qlens,
                       const int32_t* tlens, const float* logtab, int64_t logtab_len, int32_t nq,
                       int32_t mc, int32_t top_k, int32_t rank_mode, int32_t* out_target,
                       float* out_identity, float* out_rank, int32_t* out_n, int32_t* status,
                       hipStream_t st) {
  if (nq == 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<int32_t>(nq, 65536);
  topk_kernel<<<grid, TK_BLOCK, 0, st>>>(targets, counts, sims, scores, mask, qlens, tlens,
                                         logtab, logtab_len, nq, mc, top_k, rank_mode, out_target,
                                         out_identity, out_rank, out_n, status);
  return hipGetLastError();
}

}  // namespace ck
