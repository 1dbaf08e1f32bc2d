// Jaccard estimate between two sorted bottom-s sketches for each candidate
// pair (clustkit/pairwise.py:1564-1612; GPU twin pairwise.py:1664-1712).
//
// Merge walk: advance the smaller head (both on equality, counting a shared
// value); stop after s steps, when either array is exhausted, or when both
// heads are the all-ones padding.  sim = shared / steps as float32 (0 when
// no step was taken).
//
// Layout: one thread per pair walking its two sketch rows; consecutive
// values of a row share cache lines, so each 64-byte line is fetched once.
// The traffic is HBM/L2 bound (at most 2 * 8 * s bytes per pair).  The
// division is done in double and rounded once to float, as the reference's
// Numba form does.
#include "ck_internal.h"

namespace ck {

__global__ __launch_bounds__(256) void ck_jaccard_kernel(
    const int32_t* __restrict__ pairs, int64_t M, const uint64_t* __restrict__ sketches,
    int32_t s, float* __restrict__ out) {
  const uint64_t SENT = ~0ull;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < M;
       p += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t* X = sketches + (int64_t)pairs[2 * p] * s;
    const uint64_t* Y = sketches + (int64_t)pairs[2 * p + 1] * s;
    int32_t ix = 0, iy = 0, uni = 0, both = 0;
    while (uni < s && ix < s && iy < s) {
      const uint64_t u = X[ix], v = Y[iy];
      if (u == SENT && v == SENT) break;
      if (u == v) { ++both; ++ix; ++iy; }
      else if (u < v) ++ix;
      else ++iy;
      ++uni;
    }
    out[p] = uni ? (float)((double)both / (double)uni) : 0.0f;
  }
}

}  // namespace ck
