// Batched banded Smith-Waterman for gfx950: one 16-lane DPP row per pair,
// four pairs per wavefront, persistent wavefronts pulling pairs from a
// device work queue ordered by estimated cost.
//
// Semantics: clustkit/csrc/sw_align.c:25-177 (banded Gotoh local alignment,
// BLOSUM62, open -11 / extend -1, tie rules of sw_align.c:129,141,147-151,
// first row-major maximum) -- bit-exact, including the match count that
// gives the identity.
//
// Schedule.  The rows of the shorter sequence are cut into stripes of 16.
// Lane l of a segment owns row s*16 + l + 1 of stripe s and walks its band
// positions k = 0 .. W-1 (column j = i + hint - bw + k), lane l running two
// steps behind lane l-1: at step t lane l is at k = t - 2l.  Then the cell
// above (i-1, j) was produced by lane l-1 one step earlier and the diagonal
// (i-1, j-1) two steps earlier, so both arrive through one DPP row_shr:1 per
// step; the cell to the left is the lane's own previous step.  Lane 0 takes
// its "above" values from an LDS row buffer written by lane 15 of the
// previous stripe.
//
// Arithmetic.  Each DP value is one int32 key (score<<16 | pref<<13 |
// matches, see ck_internal.h) so that a single signed max() both picks the
// winning score and carries the winner's match count with the reference's
// tie preferences.  A cell costs ~20 VALU ops (2 DPP moves, 1 add, 3 maxes,
// and/or/sub normalisations, 2 for the running maximum).  Scores live in
// the 16-bit half of the key; a pair whose best score reaches SAT_LIMIT is
// handed to the int32 kernel.
//
// Out-of-matrix cells.  Rows past the shorter sequence and columns outside
// 1..lb read the padding code 31, whose substitution score is -16000.  Such
// cells never feed a cell inside the matrix (values flow right and down
// only, and columns < 1 stay at H = 0), and any value they take is bounded
// by an in-matrix value minus a gap penalty, so they can neither raise the
// maximum nor tie it.
#include "ck_internal.h"

namespace ck {

__device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

// Dynamic LDS layout per block:
//   int32  table[32*32]                       diagonal addends (score+10)<<16|pref|match
//   per segment (16 per block):
//     int2  rowbuf[rb_stride]                 (HN, EN) of the previous stripe's last row
//     uint8 bstage[bs_stride]                 residue codes of the stripe's columns
__global__ __launch_bounds__(WAVE_BLOCK) void sw_wave_kernel(
    const int32_t* __restrict__ queue, Counters* __restrict__ ctr,
    const int32_t* __restrict__ pairs, const uint8_t* __restrict__ flat,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ lengths,
    const int32_t* __restrict__ hints, int32_t band, const int8_t* __restrict__ sub,
    float threshold, uint32_t flags, int32_t rb_stride, int32_t bs_stride,
    float* __restrict__ out_sims, int32_t* __restrict__ out_scores,
    int32_t* __restrict__ out_matches, uint8_t* __restrict__ out_mask,
    int32_t* __restrict__ rerun_list) {
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* table = reinterpret_cast<int32_t*>(smem);
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  const int seg = lane / SEG;
  const int lr = lane & (SEG - 1);
  const int sidx = wv * SEGS_PER_WAVE + seg;
  int2* rowbuf = reinterpret_cast<int2*>(smem + TABLE_DIM * TABLE_DIM * 4) + sidx * rb_stride;
  uint8_t* bst = smem + TABLE_DIM * TABLE_DIM * 4 +
                 (WAVE_BLOCK / SEG) * rb_stride * (int)sizeof(int2) + sidx * bs_stride;

  for (int x = threadIdx.x; x < TABLE_DIM * TABLE_DIM; x += blockDim.x) {
    const int a = x >> 5, b = x & 31;
    int32_t s, m = 0;
    if (a == CODE_POISON || b == CODE_POISON) {
      s = POISON_SCORE;
    } else {
      s = (a < 20 && b < 20) ? (int32_t)sub[a * 20 + b] : UNKNOWN_SCORE;
      m = (a == b) ? 1 : 0;
    }
    table[x] = key_of(s + 10, PREF_D, m);
  }
  __syncthreads();

  const int32_t qlen = ctr->queue_len;
  for (;;) {
    int32_t base = 0;
    if (lane == 0) base = atomicAdd(&ctr->queue_head, SEGS_PER_WAVE);
    base = __shfl(base, 0);
    if (base >= qlen) break;
    const int32_t q = base + seg;
    const bool valid = q < qlen;
    const int64_t p = valid ? (int64_t)queue[q] : 0;
    PairGeom g;
    g.a_off = 0; g.b_off = 0; g.la = 0; g.lb = 0; g.bw = 0; g.hint = 0;
    if (valid) g = pair_geom(pairs, p, offsets, lengths, hints, band);
    const int32_t W = valid ? 2 * g.bw + 1 : 0;
    const int32_t nstr = valid ? (g.la + SEG - 1) / SEG : 0;
    int32_t wmax = W, smax = nstr;
#pragma unroll
    for (int off = SEG; off < WAVE; off <<= 1) {
      wmax = imax(wmax, __shfl_xor(wmax, off));
      smax = imax(smax, __shfl_xor(smax, off));
    }
    const int32_t T = wmax + 2 * (SEG - 1);

    for (int x = lr; x < T + 1; x += SEG) rowbuf[x] = make_int2(HN_ZERO, EN_NEG);
    int32_t seg_best = HN_ZERO;
    int32_t bad = 0;
    const uint8_t* A = flat + g.a_off;
    const uint8_t* B = flat + g.b_off;

    for (int32_t s = 0; s < smax; ++s) {
      const bool live = s < nstr;
      const int32_t Wl = live ? W : 0;
      // Residue codes of this stripe: row residue per lane, column residues
      // for positions x = j - cbase + 15 of the stripe in the LDS stage.
      const int32_t i = s * SEG + lr + 1;
      int32_t acode = CODE_POISON;
      if (live && i <= g.la) {
        const int32_t r = A[i - 1];
        bad |= r > CODE_MAX_FAST;
        acode = r > CODE_MAX_FAST ? CODE_MAX_FAST : r;
      }
      const int32_t a32 = acode * TABLE_DIM;
      const int32_t cbase = s * SEG + 1 + g.hint - g.bw;
      for (int x = lr; x < T + SEG; x += SEG) {
        const int32_t j = cbase + x - (SEG - 1);
        int32_t code = CODE_POISON;
        if (live && j >= 1 && j <= g.lb) {
          const int32_t r = B[j - 1];
          bad |= r > CODE_MAX_FAST;
          code = r > CODE_MAX_FAST ? CODE_MAX_FAST : r;
        }
        bst[x] = (uint8_t)code;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

      // Band positions of lane l run over steps t in [2l, 2l + Wl).  Steps
      // [30, wsteady) have every lane of every live segment inside its band
      // and no lane at its last position (lane 0's last position reads the
      // outside entry rowbuf[W]); they run without masks, unrolled, with the
      // LDS loads of a block issued ahead of its arithmetic.  The ramps
      // around them run the masked step.
      int32_t wlive = live ? Wl : 0x7FFFFFFF;
#pragma unroll
      for (int off = SEG; off < WAVE; off <<= 1) wlive = min(wlive, __shfl_xor(wlive, off));
      const int32_t ramp = 2 * (SEG - 1);
      const int32_t wsteady = wlive > ramp ? (wlive < T ? wlive : T) : ramp;

      int32_t hu = rowbuf[0].x;
      int32_t hout = HN_ZERO, eout = EN_NEG;
      int32_t hl = HN_ZERO, fl = FN_NEG, best = HN_ZERO;
      const uint8_t* bl = bst + (SEG - 1) - lr;
      const int32_t* tab = table + a32;
      int2* rwr = rowbuf - 2 * (SEG - 1);   // lane 15 writes position t - 30 at step t
      const bool writer = lr == SEG - 1;

      auto masked_step = [&](int32_t t) {
        const int2 rb = rowbuf[t + 1];
        const int32_t hd = hu;
        hu = __builtin_amdgcn_update_dpp(rb.x, hout, 0x111, 0xF, 0xF, false);
        const int32_t eu = __builtin_amdgcn_update_dpp(rb.y, eout, 0x111, 0xF, 0xF, false);
        const int32_t k = t - 2 * lr;
        if ((uint32_t)k < (uint32_t)Wl) {
          const int32_t d = hd + tab[bl[t]];
          int32_t e = (imax(eu, hu) | PREF_E) - KEY_ONE;
          if (k == Wl - 1) e = EN_EDGE;
          const int32_t f = (imax(fl, hl) | PREF_F) - KEY_ONE;
          const int32_t h = imax(imax(d, e), imax(f, K0_SHIFTED));
          const int32_t hn = (h & ~PREF_MASK) - 10 * KEY_ONE;
          if ((hn >> KEY_SHIFT) > (best >> KEY_SHIFT)) best = hn;
          hl = hn;
          fl = f;
          hout = hn;
          eout = e;
          if (writer) rowbuf[k] = make_int2(hn, e);
        }
      };

      int32_t t = 0;
      for (; t < ramp && t < T; ++t) masked_step(t);
      constexpr int U = 8;
      for (; t + U <= wsteady; t += U) {
        int32_t tv[U];
        int2 rb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) tv[u] = tab[bl[t + u]];
#pragma unroll
        for (int u = 0; u < U; ++u) rb[u] = rowbuf[t + u + 1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t hd = hu;
          hu = __builtin_amdgcn_update_dpp(rb[u].x, hout, 0x111, 0xF, 0xF, false);
          const int32_t eu = __builtin_amdgcn_update_dpp(rb[u].y, eout, 0x111, 0xF, 0xF, false);
          const int32_t d = hd + tv[u];
          const int32_t e = (imax(eu, hu) | PREF_E) - KEY_ONE;
          const int32_t f = (imax(fl, hl) | PREF_F) - KEY_ONE;
          const int32_t h = imax(imax(d, e), imax(f, K0_SHIFTED));
          const int32_t hn = (h & ~PREF_MASK) - 10 * KEY_ONE;
          if ((hn >> KEY_SHIFT) > (best >> KEY_SHIFT)) best = hn;
          hl = hn;
          fl = f;
          hout = hn;
          eout = e;
          if (writer) rwr[t + u] = make_int2(hn, e);
        }
      }
      for (; t < T; ++t) masked_step(t);
      // First maximum of the stripe: highest score, then lowest row (lane).
      int32_t r = (int32_t)((uint32_t)best & 0xFFFF0000u) | (SEG - 1 - lr);
#pragma unroll
      for (int off = SEG / 2; off >= 1; off >>= 1) r = imax(r, __shfl_xor(r, off));
      const int32_t wl = (SEG - 1) - (r & (SEG - 1));
      const int32_t wbest = __shfl(best, (lane & ~(SEG - 1)) | wl);
      if (live && (wbest >> KEY_SHIFT) > (seg_best >> KEY_SHIFT)) seg_best = wbest;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    int32_t any_bad = bad;
#pragma unroll
    for (int off = 1; off < SEG; off <<= 1) any_bad |= __shfl_xor(any_bad, off);
    if (valid && lr == 0) {
      int32_t score = (seg_best >> KEY_SHIFT) + 11;
      int32_t matches = seg_best & MATCH_MASK;
      if (any_bad || score >= SAT_LIMIT) {
        const int32_t slot = atomicAdd(&ctr->rerun_len, 1);
        rerun_list[slot] = (int32_t)p;
      } else {
        float ident = 0.0f;
        if (score > 0) ident = (float)matches / (float)g.la;
        else { score = 0; matches = 0; }
        out_sims[p] = ident;
        out_scores[p] = score;
        if (out_matches) out_matches[p] = matches;
        if (out_mask)
          out_mask[p] = (flags & 0x2u) ? (uint8_t)(score > 0 && ident >= threshold)
                                       : (uint8_t)(score > 0);
      }
    }
  }
}

}  // namespace ck
