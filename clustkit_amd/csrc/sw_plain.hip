// Plain banded Smith-Waterman: one pair per thread, int32 arithmetic.
//
// This is the straightforward GPU form of clustkit/csrc/sw_align.c:25-177
// (and of the reference's one-thread-per-pair GPU kernel,
// clustkit/pairwise.py:1215-1402): each thread walks its pair row by row.
// It serves three roles: the "plain" form the batched kernel is measured
// against, the exact int32 re-run for pairs the batched kernel flags
// (saturated 16-bit scores, shorter sequence > 8191, very wide bands, raw
// residue codes > 30), and the general path for any input.
//
// Workspace: ONE band row per thread, band-relative, updated in place.
// Row i's band position c maps to column j = c + (i + hint - bw).  The
// previous row's cell at the same position c is column j - 1 (diagonal) and
// position c + 1 is column j (above), so processing c upward and writing
// position c after reading c + 1 needs a single array of 2*bw + 2 entries.
// Entries never written hold "outside the band" (H = 0, Hm = 0, E = -inf).
// The row is interleaved across threads (entry c of thread t at c*T + t) so
// a wavefront's accesses coalesce into 16-byte-per-lane transactions.
#include "ck_internal.h"

namespace ck {

__global__ __launch_bounds__(256) void sw_plain_kernel(
    const int32_t* __restrict__ pairs, const int32_t* __restrict__ list,
    const int32_t* __restrict__ list_len_dev, int64_t list_len_host, int64_t list_base,
    int64_t batch_cap,
    const uint8_t* __restrict__ flat, const int64_t* __restrict__ offsets,
    const int32_t* __restrict__ lengths, const int32_t* __restrict__ hints, int32_t band,
    const int8_t* __restrict__ sub, float threshold, uint32_t flags, int4* __restrict__ ws,
    int32_t ws_width, float* __restrict__ out_sims, int32_t* __restrict__ out_scores,
    int32_t* __restrict__ out_matches, uint8_t* __restrict__ out_mask,
    int32_t* __restrict__ err) {
  __shared__ int8_t s_sub[400];
  for (int x = threadIdx.x; x < 400; x += blockDim.x) s_sub[x] = sub[x];
  __syncthreads();

  int64_t n = (list_len_dev ? (int64_t)(*list_len_dev) : list_len_host) - list_base;
  if (n > batch_cap) n = batch_cap;
  const int64_t T = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int4* row = ws + tid;

  for (int64_t w = tid; w < n; w += T) {
    const int64_t p = list ? (int64_t)list[list_base + w] : list_base + w;
    const int32_t x = pairs[2 * p], y = pairs[2 * p + 1];
    const int32_t lx = lengths[x], ly = lengths[y];
    int32_t score = 0, matches = 0;
    float ident = 0.0f;
    bool skip = (lx == 0 || ly == 0) ||
                ((flags & 0x1u) && length_ratio_fails(lx, ly, threshold));
    if (!skip) {
      PairGeom g = pair_geom(pairs, p, offsets, lengths, hints, band);
      const uint8_t* A = flat + g.a_off;
      const uint8_t* B = flat + g.b_off;
      const int32_t bw = g.bw;
      const int32_t width = 2 * bw + 2;
      if (width <= ws_width) {
        for (int32_t c = 0; c < width; ++c) row[(int64_t)c * T] = make_int4(0, 0, NEG_INF, 0);
        for (int32_t i = 1; i <= g.la; ++i) {
          const int32_t c0 = i + g.hint - bw;          // column of band position 0
          int32_t lo = c0 < 1 ? 1 : c0;
          int32_t hi = c0 + 2 * bw;
          if (hi > g.lb) hi = g.lb;
          if (lo > hi) {
            if (c0 > g.lb) break;                        // band has left the matrix
            continue;                                    // band not yet inside
          }
          const uint8_t ra = A[i - 1];
          const int8_t* srow = s_sub + (ra < 20 ? ra * 20 : 0);
          int32_t f = NEG_INF, fm = 0, hl = 0, hlm = 0;
          int4 dg = row[(int64_t)(lo - c0) * T];         // (i-1, lo-1)
          for (int32_t j = lo; j <= hi; ++j) {
            const int32_t c = j - c0;
            const int4 up = row[(int64_t)(c + 1) * T];    // (i-1, j)
            const uint8_t rb = B[j - 1];
            const int32_t s = (ra < 20 && rb < 20) ? (int32_t)srow[rb] : UNKNOWN_SCORE;
            const int32_t dv = dg.x + s, dm = dg.y + (ra == rb ? 1 : 0);
            int32_t ev, em;
            if (up.z + GAP_EXTEND >= up.x + GAP_OPEN) { ev = up.z + GAP_EXTEND; em = up.w; }
            else { ev = up.x + GAP_OPEN; em = up.y; }
            const int32_t fx = (j == lo) ? NEG_INF : f + GAP_EXTEND;
            if (!(fx >= hl + GAP_OPEN)) { f = hl + GAP_OPEN; fm = hlm; } else { f = fx; }
            int32_t h = 0, hm = 0;
            if (dv > h) { h = dv; hm = dm; }
            if (ev > h) { h = ev; hm = em; }
            if (f > h) { h = f; hm = fm; }
            row[(int64_t)c * T] = make_int4(h, hm, ev, em);
            hl = h; hlm = hm;
            if (h > score) { score = h; matches = hm; }
            dg = up;
          }
        }
        if (score > 0) ident = (float)matches / (float)g.la;
        else { score = 0; matches = 0; }
      } else {
        // Band row wider than the workspace (the caller's max_len is below
        // this pair's longer sequence): zero outputs and the call's error
        // flag, which the host entry point turns into CK_ERR_INVALID.
        atomicOr(err, 1);
      }
    }
    out_sims[p] = ident;
    out_scores[p] = score;
    if (out_matches) out_matches[p] = matches;
    if (out_mask)
      out_mask[p] = (flags & 0x2u) ? (uint8_t)(score > 0 && ident >= threshold)
                                   : (uint8_t)(score > 0);
  }
}

}  // namespace ck
