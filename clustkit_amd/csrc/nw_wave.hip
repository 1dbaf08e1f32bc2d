// Batched banded global alignment (Needleman-Wunsch identity) for gfx950.
//
// Semantics: clustkit/pairwise.py:258-362 (_nw_identity: +2 / -1 on byte
// equality, gaps -10 / -1) and 365-505 (_nw_identity_and_score: substitution
// table for codes < 20, -4 otherwise, gaps -11 / -1), with the batch rules of
// pairwise.py:684-736 / 804-850: length-ratio test, length-difference test,
// adaptive half band min(bw, max(10, diff + 10)), the shorter sequence on
// rows, full matrix when the band is <= 0 or the longer sequence has at most
// 50 residues, early stop after any row whose best match count cannot reach
// the threshold, mask = identity >= threshold.  Bit-exact: identity, score
// (scored form) and mask.
//
// Cell set (equivalent to the reference's toggled rows with border patches):
// row 0 holds columns 0..min(bw, lb) (H = 0 at column 0, open + ext*(j-1)),
// column 0 holds rows 1..bw+1 (open + ext*(i-1)), and inner cells satisfy
// |i - j| <= bw.  Anything else reads as H = E = -1,000,000 with zero match
// counts; values derived from that sentinel never win a comparison against a
// value of the cell set, because every inner cell has a diagonal candidate
// from the cell set.
//
// Band positions.  Row i position k (0 <= k < W = 2bw+1) is column
// j = i - bw + k; the cell above (i-1, j) is position k+1 of row i-1, the
// diagonal (i-1, j-1) position k.
//
// Kernels:
//   nw_pack_kernel     S = 8 lanes per pair (windows up to 63 positions, two
//                      pairs per DPP row) or 16 (wider windows), stripes of S
//                      rows; lane l owns row S*s + l + 1 and runs two steps
//                      behind lane l-1, so the cell above arrives from lane
//                      l-1 one step late and the diagonal two steps late
//                      through DPP row_shr:1; lane 0 reads them from an LDS
//                      row buffer written by the previous stripe.  Values are
//                      packed keys (see below); pairs whose scores fit 16
//                      bits.  Persistent wavefronts take pairs from the
//                      8-lane range of a queue sorted by cost, then from the
//                      16-lane range.
//   nw_wave_kernel<16> the same schedule with int32 values, for pairs whose
//                      scores may leave 16 bits.
//   nw_plain_kernel    one pair per thread over one band row in global memory
//                      (updated in place), for bands wider than the LDS row
//                      buffer.
#include <type_traits>

#include "ck_internal.h"

namespace ck {

namespace {

constexpr int32_t NW_NEG = -1000000;
constexpr int NW_TAB = 21;              // codes >= 20 share row/column 20 (score -4)
constexpr int NW_TAB_WORDS = 448;       // 21 * 21 rounded up
constexpr int NW_SB_PAD = 48;           // staged codes beyond the band: W + 3S - 1 used

struct NwGeom {
  int64_t a_off, b_off;   // a = rows (shorter), b = columns
  int32_t la, lb, bw, W;
  int32_t status;         // 0 align, 1 filtered (identity 0, score NEG), 2 empty input
};

// Batch rules of pairwise.py:694-736 / 812-850 and the band rule of
// pairwise.py:285-287.
// raw: the inner function alone (pairwise.py:258-287): no length tests, the
// band as given.
__device__ __forceinline__ NwGeom nw_geom(const int32_t* pairs, int64_t p, const int64_t* offsets,
                                          const int32_t* lengths, int32_t band_width,
                                          float threshold, int32_t raw) {
  NwGeom g;
  const int32_t x = pairs[2 * p], y = pairs[2 * p + 1];
  const int32_t lx = lengths[x], ly = lengths[y];
  const int32_t sh = lx < ly ? lx : ly, lg = lx < ly ? ly : lx;
  const int32_t diff = lg - sh;
  g.status = 0;
  if (!raw && ((lg > 0 && (float)sh / (float)lg < threshold) || diff > band_width)) g.status = 1;
  int32_t ab = diff + 10 > 10 ? diff + 10 : 10;
  if (band_width < ab) ab = band_width;
  if (raw) ab = band_width;
  // Shorter sequence on rows; equal lengths keep the pair's order.
  const bool keep = lx <= ly;
  g.a_off = offsets[keep ? x : y];
  g.b_off = offsets[keep ? y : x];
  g.la = sh;
  g.lb = lg;
  if (g.status == 0 && sh == 0) g.status = 2;
  g.bw = (ab <= 0 || lg <= 50) ? lg : ab;
  g.W = 2 * g.bw + 1;
  return g;
}

__device__ __forceinline__ void nw_write(int64_t p, float sim, int32_t score, float threshold,
                                         float* out_sims, int32_t* out_scores,
                                         uint8_t* out_mask) {
  out_sims[p] = sim;
  if (out_scores) out_scores[p] = score;
  out_mask[p] = sim >= threshold ? 1 : 0;
}

// Row-0 entry of band position k: column j = k - bw.
__device__ __forceinline__ int4 nw_row0(int32_t k, int32_t bw, int32_t lb, int32_t go) {
  const int32_t j = k - bw;
  int32_t h = NW_NEG;
  if (j == 0) h = 0;
  else if (j >= 1 && j <= bw && j <= lb) h = go + (-1) * (j - 1);
  return make_int4(h, 0, NW_NEG, 0);
}

// LDS bytes of one 16-lane segment: row buffer + staging, 16-byte aligned.
__host__ __device__ __forceinline__ int32_t nw_seg_bytes(int32_t w_cap) {
  return (w_cap + 1) * 16 + ((w_cap + NW_SB_PAD + 3) & ~3) * 4;
}

}  // namespace

// Per pair: answer filtered and empty pairs, send pairs whose band row does
// not fit the LDS buffer to the one-pair-per-thread list, give the rest a
// bin in the range of its kernel: cost level (stripes x (W + 2(S - 1))
// steps, descending) and window width.
__global__ __launch_bounds__(256) void nw_prep_kernel(
    const int32_t* __restrict__ pairs, int64_t M, const int64_t* __restrict__ offsets,
    const int32_t* __restrict__ lengths, int32_t band_width, float threshold, int32_t raw,
    int32_t w_cap, int32_t w_narrow, int32_t smax, int32_t smin, int32_t* __restrict__ bin_of,
    int32_t* __restrict__ hist,
    Counters* __restrict__ ctr, int32_t* __restrict__ plain_list, float* __restrict__ out_sims,
    int32_t* __restrict__ out_scores, uint8_t* __restrict__ out_mask) {
  __shared__ int32_t lhist[NBINS];
  for (int x = threadIdx.x; x < NBINS; x += blockDim.x) lhist[x] = 0;
  __syncthreads();
  int32_t skipped = 0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < M;
       p += (int64_t)gridDim.x * blockDim.x) {
    const NwGeom g = nw_geom(pairs, p, offsets, lengths, band_width, threshold, raw);
    if (g.status) {
      // Removed by a length test: mask 0 whatever the threshold
      // (pairwise.py:716-726); empty input: identity 0, mask 0 >= t.
      nw_write(p, 0.0f, g.status == 1 ? NW_NEG : 0, g.status == 1 ? 2.0f : threshold, out_sims,
               out_scores, out_mask);
      bin_of[p] = -1;
      skipped += g.status == 1;
      continue;
    }
    if (g.W > w_cap) {
      bin_of[p] = -1;
      const int32_t slot = atomicAdd(&ctr->plain_len, 1);
      atomicAdd(&ctr->prep_plain, 1);
      plain_list[slot] = (int32_t)p;
      continue;
    }
    // Kernel class, each with NBINS/4 bins in descending cost:
    //   0  nw_wave_kernel<16>  scores may leave 16 bits
    //   1  nw_pack_kernel, 16 lanes   packed keys, W > w_narrow
    //   2  nw_pack_kernel, 8 lanes    packed keys, W <= w_narrow
    // Packed keys need la <= 8191 (13-bit match counts), every cell score
    // at most smax * la, plus the pipelined kernel's diagonal offset of up
    // to la + lb + 1 (<= 32000), and at least smin * la - 11 - lb, kept
    // above -31000 so the -32000 sentinel stays below every real value.
    const bool pack = g.la <= 8191 && (int64_t)smax * g.la + g.la + g.lb <= 32000 &&
                      (int64_t)(-smin) * g.la + g.lb <= 31000;
    const int32_t cls = !pack ? 0 : (g.W <= w_narrow ? 2 : 1);
    const int32_t S = cls == 2 ? 8 : 16;
    // 64 cost levels, eight per doubling of the step count (2^7 .. 2^15
    // steps), each split into 16 sub-bins by window width over the class's
    // width range: the segments of one wavefront run in lockstep, so pairs
    // that share a wavefront should agree in window (steps per stripe) and
    // stripe count, not only in their product.
    const uint32_t steps = (uint32_t)min<int64_t>(
        (int64_t)((g.la + S - 1) / S) * (g.W + 2 * (S - 1)), 0x7FFFFFFF);
    const int32_t e = 31 - __clz((int32_t)max(steps, 1u));
    const int32_t fr = e >= 3 ? (int32_t)(steps >> (e - 3)) & 7 : 0;
    const int32_t lv = min(max(e * 8 + fr - 56, 0), 63);
    const int32_t wlo = cls == 1 ? w_narrow + 1 : 1;
    const int32_t whi = cls == 2 ? w_narrow : w_cap;
    const int32_t sub =
        min(max((int32_t)((int64_t)(g.W - wlo) * 16 / max(whi - wlo + 1, 1)), 0), 15);
    const int32_t bin = (cls + 1) * (NBINS / 4) - 1 - (lv * 16 + sub);
    bin_of[p] = bin;
    atomicAdd(&lhist[bin], 1);
  }
  __syncthreads();
  for (int x = threadIdx.x; x < NBINS; x += blockDim.x)
    if (lhist[x]) atomicAdd(&hist[x], lhist[x]);
  if (skipped) atomicAdd(&ctr->skipped, skipped);
}

// LDS layout (dynamic): [table: NW_TAB_WORDS ints] then per S-lane segment
// a row buffer of w_cap + 1 int4 (H, matches, E, E matches) and a staging
// buffer of w_cap + NW_SB_PAD ints (raw code | min(code, 20) << 8).
// S = 16: one pair per DPP row; S = 8: two pairs per DPP row, the second
// pair's first lane taking the row buffer instead of its DPP neighbour.
// The wavefronts take queue slots [*q_begin, *q_end) through *head.
template <bool SCORED, int S>
__global__ __launch_bounds__(WAVE_BLOCK) void nw_wave_kernel(
    const int32_t* __restrict__ queue, const int32_t* __restrict__ q_begin,
    const int32_t* __restrict__ q_end, int32_t* __restrict__ head,
    const int32_t* __restrict__ pairs, const uint8_t* __restrict__ flat,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ lengths,
    int32_t band_width, float threshold, int32_t raw, const int8_t* __restrict__ sub,
    int32_t w_cap, float* __restrict__ out_sims, int32_t* __restrict__ out_scores,
    uint8_t* __restrict__ out_mask) {
  static_assert(S == 8 || S == 16, "8 or 16 lanes per pair");
  constexpr int PER_WAVE = WAVE / S;
  constexpr uint64_t SEG_BITS = (1ull << S) - 1;
  extern __shared__ int4 nw_lds[];
  int32_t* tab = reinterpret_cast<int32_t*>(nw_lds);
  const int lane = threadIdx.x & (WAVE - 1);
  const int l = lane & (S - 1);
  const int seg = lane / S;
  const int sg = (threadIdx.x >> 6) * PER_WAVE + seg;
  char* seg_base = reinterpret_cast<char*>(nw_lds) + NW_TAB_WORDS * 4 +
                   (size_t)sg * nw_seg_bytes(w_cap);
  int4* rb = reinterpret_cast<int4*>(seg_base);
  int32_t* sb = reinterpret_cast<int32_t*>(seg_base + (size_t)(w_cap + 1) * sizeof(int4));
  if (SCORED) {
    for (int x = threadIdx.x; x < NW_TAB * NW_TAB; x += blockDim.x) {
      const int r = x / NW_TAB, c = x % NW_TAB;
      tab[x] = (r < 20 && c < 20) ? (int32_t)sub[r * 20 + c] : -4;
    }
  }
  __syncthreads();
  constexpr int32_t GO = SCORED ? -11 : -10;
  constexpr int32_t GE = -1;
  const int32_t qb = q_begin ? *q_begin : 0;
  const int32_t n = *q_end;
  for (;;) {
    const int32_t q0 = qb + wave_fetch(head, PER_WAVE);
    if (q0 >= n) break;
    const int32_t slot = q0 + seg;
    if (slot >= n) continue;
    const int64_t p = queue[slot];
    const NwGeom g = nw_geom(pairs, p, offsets, lengths, band_width, threshold, raw);
    const uint8_t* A = flat + g.a_off;
    const uint8_t* B = flat + g.b_off;
    const int32_t la = g.la, lb = g.lb, bw = g.bw, W = g.W;
    // Row 0.
    for (int32_t k = l; k <= W; k += S) rb[k] = nw_row0(k, bw, lb, GO);
    __builtin_amdgcn_wave_barrier();
    bool stopped = false;
    int32_t fh = NW_NEG, fm = 0;
    const int32_t nstripes = (la + S - 1) / S;
    for (int32_t s = 0; s < nstripes; ++s) {
      const int32_t i0 = S * s + 1;
      const int32_t i = i0 + l;
      const bool real = i <= la;
      // Stage columns jb .. jb + W + 3S - 2 of b (jb = i0 - bw - (S - 1)).
      const int32_t jb = i0 - bw - (S - 1);
      for (int32_t x = l; x < W + 3 * S - 1; x += S) {
        const int32_t j = jb + x;
        const uint32_t c = (j >= 1 && j <= lb) ? B[j - 1] : 0xFFu;
        sb[x] = (int32_t)(c | ((c < 20u ? c : 20u) << 8));
      }
      __builtin_amdgcn_wave_barrier();
      const uint32_t xa = real ? A[i - 1] : 0xFEu;
      const int32_t rowbase = SCORED ? (int32_t)(xa < 20u ? xa : 20u) * NW_TAB : 0;
      // Column 0 of this row: position kr (or just before the first
      // position), H = open + ext*(i-1) for i <= bw + 1.
      const int32_t kr = i <= bw ? bw - i : -1;
      const int32_t hr = i <= bw + 1 ? GO + GE * (i - 1) : NW_NEG;
      // Band positions of this row's cells (rows past la: any range, they
      // never stop the pair).
      const int32_t klo = bw + 1 - i > 0 ? bw + 1 - i : 0;
      const uint32_t kspan = (uint32_t)((lb - i + bw < W - 1 ? lb - i + bw : W - 1) - klo);
      // The lane that writes the row buffer: lane S-1, or in the last
      // stripe the lane of row la, whose entry at lb - la + bw is then the
      // result.
      const int32_t wl = la - i0 < S - 1 ? la - i0 : S - 1;
      int32_t hout = hr, mout = 0, eout = NW_NEG, emout = 0;
      int32_t F = NW_NEG, FM = 0, rmax = 0;
      const int4 r0 = rb[0];
      int32_t hu = r0.x, mu = r0.y;
      const int32_t T = W + 2 * (S - 1);
      // Loads run one step ahead: the row-buffer entry for lane 0 (read by
      // every lane of the segment from the same address) and the code.
      int4 rn = rb[min(1, W)];
      int32_t vn = sb[S - 1 - l];
      auto step = [&](int32_t t) {
        const int32_t k = t - 2 * l;
        const int4 r = rn;
        const int32_t v = vn;
        rn = rb[min(t + 2, W)];
        vn = sb[t + S - l];
        const int32_t hd = hu, md = mu;
        hu = __builtin_amdgcn_update_dpp(r.x, hout, 0x111, 0xF, 0xF, false);
        mu = __builtin_amdgcn_update_dpp(r.y, mout, 0x111, 0xF, 0xF, false);
        int32_t eu = __builtin_amdgcn_update_dpp(r.z, eout, 0x111, 0xF, 0xF, false);
        int32_t emu = __builtin_amdgcn_update_dpp(r.w, emout, 0x111, 0xF, 0xF, false);
        if (S == 8 && l == 0) {   // second pair of the DPP row: not lane 7's cell
          hu = r.x;
          mu = r.y;
          eu = r.z;
          emu = r.w;
        }
        const int32_t mt = (uint32_t)(v & 0xFF) == xa ? 1 : 0;
        const int32_t sc = SCORED ? tab[rowbase + (v >> 8)] : (mt ? 2 : -1);
        const int32_t d = hd + sc, dm = md + mt;
        const int32_t eext = eu + GE, eop = hu + GO;
        int32_t e = eext >= eop ? eext : eop;
        const int32_t em = eext >= eop ? emu : mu;
        if (k == W - 1) e = NW_NEG;            // the cell above is outside the band
        const int32_t fext = F + GE, fop = hout + GO;
        if (fext >= fop) {
          F = fext;
        } else {
          F = fop;
          FM = mout;
        }
        int32_t h = d, m = dm;
        if (e > h) { h = e; m = em; }
        if (F > h) { h = F; m = FM; }
        if (k == kr) { h = hr; m = 0; F = NW_NEG; }
        hout = h;
        mout = m;
        eout = e;
        emout = em;
        if ((uint32_t)(k - klo) <= kspan) rmax = m > rmax ? m : rmax;
        if (l == wl && (uint32_t)k < (uint32_t)W) rb[k] = make_int4(h, m, e, em);
      };
      int32_t t = 0;
      for (; t + 1 < T; t += 2) {
        step(t);
        step(t + 1);
      }
      if (t < T) step(t);
      __builtin_amdgcn_wave_barrier();
      // Early stop after any row whose best match count cannot reach t.
      const bool stop_row = real && (float)(rmax + (la - i)) / (float)la < threshold;
      const uint64_t bal = __ballot(stop_row);
      if ((bal >> (seg * S)) & SEG_BITS) {
        stopped = true;
        break;
      }
    }
    if (!stopped && lb - la <= bw) {
      const int4 f = rb[lb - la + bw];
      fh = f.x;
      fm = f.y;
    }
    if (l == 0) {
      if (stopped) nw_write(p, 0.0f, NW_NEG, threshold, out_sims, out_scores, out_mask);
      else nw_write(p, (float)fm / (float)la, fh, threshold, out_sims, out_scores, out_mask);
    }
  }
}

// ---------------------------------------------------------------------------
// Packed-key form.  Each DP value is one int32 key score << 16 | pref << 13 |
// matches, so one signed max() picks the winner with the reference's tie
// rules and carries its match count:
//   E = extension (pref PE) over opening (pref 0), F likewise with PF;
//   H = diagonal (PD) over E (PE) over F (PF).
// Stored values are the candidates the neighbours need:
//   HN = (H + open) << 16 | m             gap opening, pref 0
//   EN = (E + ext) << 16 | PE | em         vertical extension
//   FN = (F + ext) << 16 | PF | fm         horizontal extension
// and the diagonal candidate is HN of the diagonal cell plus an addend
// ((s - open) << 16 | PD) + match.  Lanes exchange HN and EN only.  Usable
// when every score of the pair's cell set lies inside 16 bits with room for
// the sentinel (the prep kernel checks la <= 8191 and the score bounds;
// other pairs run nw_wave_kernel).
// ---------------------------------------------------------------------------

namespace {

constexpr int32_t NWK_ONE = 1 << 16;
constexpr int32_t NWK_PD = 3 << 13;
constexpr int32_t NWK_PE = 2 << 13;
constexpr int32_t NWK_PF = 1 << 13;
constexpr int32_t NWK_PMASK = 3 << 13;
constexpr int32_t NWK_MMASK = (1 << 13) - 1;
constexpr int32_t NWK_NEG = -32000;          // "-inf" score inside keys

__host__ __device__ constexpr int32_t nwk(int32_t score, int32_t pref, int32_t m) {
  return (int32_t)((uint32_t)score << 16) + pref + m;
}

// Minimum of v (0 <= v < 2^11) over the active lanes of the wavefront (one
// ballot per bit), uniform across them: segments of a wavefront run the same
// instruction stream on their own pairs.
__device__ __forceinline__ int32_t nw_active_min(int32_t v) {
  int32_t lo = 0;
#pragma unroll
  for (int b = 10; b >= 0; --b) {
    const int32_t cand = lo | (1 << b);
    if (__ballot(v < cand) == 0) lo = cand;
  }
  return lo;
}

// LDS bytes of one segment of the packed kernel: row buffer of int2
// (HN, EN) and the staged codes, one byte each (16-byte aligned).
__host__ __device__ __forceinline__ int32_t nwk_seg_bytes(int32_t w_cap) {
  return (w_cap + 1) * 8 + ((w_cap + NW_SB_PAD + 15) & ~15);
}

// LDS bytes of one wavefront's region in the packed kernel: four 16-lane
// segments of the wide class or eight 8-lane segments of the narrow one.
__host__ __device__ __forceinline__ int32_t nwk_wave_bytes(int32_t w_cap16, int32_t w_cap8) {
  const int32_t a = 4 * nwk_seg_bytes(w_cap16), b = 8 * nwk_seg_bytes(w_cap8);
  return a > b ? a : b;
}

// Row-0 entry of band position k in key form.
__device__ __forceinline__ int2 nwk_row0(int32_t k, int32_t bw, int32_t lb, int32_t go) {
  const int32_t j = k - bw;
  int32_t h = NWK_NEG;
  if (j == 0) h = 0;
  else if (j >= 1 && j <= bw && j <= lb) h = go - (j - 1);
  return make_int2(nwk(h + go, 0, 0), nwk(NWK_NEG - 1, NWK_PE, 0));
}

// The same in the pipelined kernel's diagonal-offset keys: a real HN of
// cell (0, j) carries j + 1 (the domain of the cells that read it).
__device__ __forceinline__ int2 nwp_row0(int32_t k, int32_t bw, int32_t lb, int32_t go) {
  const int32_t j = k - bw;
  int32_t h = NWK_NEG;
  if (j == 0) h = 1;
  else if (j >= 1 && j <= bw && j <= lb) h = go - (j - 1) + j + 1;
  return make_int2(nwk(h + go, 0, 0), nwk(NWK_NEG - 1, NWK_PE, 0));
}

}  // namespace

// The pairs of queue slots [qb, n) with S lanes per pair, taken PER_WAVE at
// a time through *head.  wave_lds: this wavefront's own LDS region (its
// segments' row buffers and staged codes); tab: the packed substitution
// table (scored form).
template <bool SCORED, int S>
__device__ __forceinline__ void nw_pack_pairs(
    const int32_t* __restrict__ queue, int32_t qb, int32_t n, int32_t* __restrict__ head,
    const int32_t* __restrict__ pairs, const uint8_t* __restrict__ flat,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ lengths,
    int32_t band_width, float threshold, int32_t raw, const int32_t* __restrict__ tab,
    char* wave_lds, int32_t w_cap, float* __restrict__ out_sims,
    int32_t* __restrict__ out_scores, uint8_t* __restrict__ out_mask) {
  static_assert(S == 8 || S == 16, "8 or 16 lanes per pair");
  constexpr int PER_WAVE = WAVE / S;
  constexpr int32_t GO = SCORED ? -11 : -10;
  constexpr int32_t GE = -1;
  // Lane l-1 of a pair sits one lane down (S = 16, one pair per DPP row) or
  // two (S = 8: the two pairs of a DPP row interleave, even and odd lanes);
  // each pair's lane 0 then falls off the row and keeps the row-buffer value.
  constexpr int DPP = S == 16 ? 0x111 : 0x112;   // row_shr:1 / row_shr:2
  // Diagonal addends of the plain form (match +2, mismatch -1).
  constexpr int32_t ADD_MATCH = nwk(2 - GO, NWK_PD, 1);
  constexpr int32_t ADD_MISS = nwk(-1 - GO, NWK_PD, 0);
  constexpr int32_t EK_EDGE = nwk(NWK_NEG, NWK_PE, 0);      // E from outside the band
  constexpr int32_t FN_RESET = nwk(NWK_NEG - 1, NWK_PF, 0); // no horizontal extension
  const int lane = threadIdx.x & (WAVE - 1);
  const int l = S == 16 ? (lane & 15) : ((lane & 15) >> 1);
  const int seg = S == 16 ? (lane >> 4) : ((lane >> 4) * 2 + (lane & 1));
  const uint64_t seg_lanes = S == 16 ? (0xFFFFull << (lane & ~15))
                                     : (0x5555ull << ((lane & ~15) | (lane & 1)));
  char* seg_base = wave_lds + (size_t)seg * nwk_seg_bytes(w_cap);
  int2* rb = reinterpret_cast<int2*>(seg_base);
  uint8_t* sb = reinterpret_cast<uint8_t*>(seg_base + (size_t)(w_cap + 1) * sizeof(int2));
  for (;;) {
    const int32_t q0 = qb + wave_fetch(head, PER_WAVE);
    if (q0 >= n) break;
    const int32_t slot = q0 + seg;
    if (slot >= n) continue;
    const int64_t p = queue[slot];
    const NwGeom g = nw_geom(pairs, p, offsets, lengths, band_width, threshold, raw);
    const uint8_t* A = flat + g.a_off;
    const uint8_t* B = flat + g.b_off;
    const int32_t la = g.la, lb = g.lb, bw = g.bw, W = g.W;
    for (int32_t k = l; k <= W; k += S) rb[k] = nwk_row0(k, bw, lb, GO);
    __builtin_amdgcn_wave_barrier();
    bool stopped = false;
    const int32_t nstripes = (la + S - 1) / S;
    for (int32_t s = 0; s < nstripes; ++s) {
      const int32_t i0 = S * s + 1;
      const int32_t i = i0 + l;
      const bool real = i <= la;
      const int32_t jb = i0 - bw - (S - 1);
      for (int32_t x = l; x < W + 3 * S - 1; x += S) {
        const int32_t j = jb + x;
        sb[x] = (j >= 1 && j <= lb) ? B[j - 1] : (uint8_t)0xFF;
      }
      __builtin_amdgcn_wave_barrier();
      const uint32_t xa = real ? A[i - 1] : 0xFEu;
      const int32_t rowbase = SCORED ? (int32_t)(xa < 20u ? xa : 20u) * NW_TAB : 0;
      // Column 0 of this row at position kr (or just before the first
      // position): HN of H = open + ext*(i-1) for i <= bw + 1, else -inf.
      const int32_t kr = i <= bw ? bw - i : -1;
      const int32_t hn_r = nwk((i <= bw + 1 ? GO + GE * (i - 1) : NWK_NEG) + GO, 0, 0);
      const int32_t klo = bw + 1 - i > 0 ? bw + 1 - i : 0;
      const uint32_t kspan = (uint32_t)((lb - i + bw < W - 1 ? lb - i + bw : W - 1) - klo);
      const int32_t wl = la - i0 < S - 1 ? la - i0 : S - 1;
      int32_t hout = hn_r, eout = nwk(NWK_NEG - 1, NWK_PE, 0), fn = FN_RESET, rmax = 0;
      int32_t hu = rb[0].x;
      const int32_t T = W + 2 * (S - 1);
      int2 rn = rb[min(1, W)];
      uint32_t vn = sb[S - 1 - l];
      auto step = [&](int32_t t) {
        const int32_t k = t - 2 * l;
        const int2 r = rn;
        const uint32_t v = vn;
        rn = rb[min(t + 2, W)];
        vn = sb[t + S - l];
        const int32_t hd = hu;
        hu = __builtin_amdgcn_update_dpp(r.x, hout, DPP, 0xF, 0xF, false);
        const int32_t eu = __builtin_amdgcn_update_dpp(r.y, eout, DPP, 0xF, 0xF, false);
        const int32_t mt = v == xa ? 1 : 0;
        const int32_t add = SCORED ? tab[rowbase + (int32_t)(v < 20u ? v : 20u)] + mt
                                   : (mt ? ADD_MATCH : ADD_MISS);
        const int32_t d = hd + add;
        int32_t ek = max(eu, hu) | NWK_PE;
        if (k == W - 1) ek = EK_EDGE;            // the cell above is outside the band
        const int32_t fk = max(fn, hout) | NWK_PF;
        const int32_t h = max(max(d, ek), fk);
        hout = (h & ~NWK_PMASK) + GO * NWK_ONE;
        eout = ek - NWK_ONE;
        fn = fk - NWK_ONE;
        if (k == kr) {
          hout = hn_r;
          fn = FN_RESET;
        }
        if ((uint32_t)(k - klo) <= kspan) rmax = max(rmax, h & NWK_MMASK);
        if (l == wl && (uint32_t)k < (uint32_t)W) rb[k] = make_int2(hout, eout);
      };
      // Steps at which no lane of any active segment meets a rule: every
      // lane past its first position and before the band's last one, in a
      // stripe whose real rows all lie inside the matrix region of the band
      // (i > bw: no column-0 cell; i <= lb - bw: no column past lb), so the
      // row maximum takes every cell and the row-buffer write needs no test.
      auto steady = [&](int32_t t) -> int32_t {
        const int32_t k = t - 2 * l;
        const int2 r = rn;
        const uint32_t v = vn;
        rn = rb[t + 2];
        vn = sb[t + S - l];
        const int32_t hd = hu;
        hu = __builtin_amdgcn_update_dpp(r.x, hout, DPP, 0xF, 0xF, false);
        const int32_t eu = __builtin_amdgcn_update_dpp(r.y, eout, DPP, 0xF, 0xF, false);
        const int32_t mt = v == xa ? 1 : 0;
        const int32_t add = SCORED ? tab[rowbase + (int32_t)(v < 20u ? v : 20u)] + mt
                                   : (mt ? ADD_MATCH : ADD_MISS);
        const int32_t d = hd + add;
        const int32_t ek = max(eu, hu) | NWK_PE;
        const int32_t fk = max(fn, hout) | NWK_PF;
        const int32_t h = max(max(d, ek), fk);
        hout = (h & ~NWK_PMASK) + GO * NWK_ONE;
        eout = ek - NWK_ONE;
        fn = fk - NWK_ONE;
        if (l == wl) rb[k] = make_int2(hout, eout);
        return h & NWK_MMASK;
      };
      const int32_t ilast = min(i0 + S - 1, la);
      const bool inner = i0 > bw && ilast <= lb - bw;
      const int32_t t_lo = 2 * (S - 1);
      const int32_t t_hi = nw_active_min(inner ? W - 1 : 0);
      int32_t t = 0;
      if (t_hi > t_lo) {
#pragma unroll 1
        for (; t < t_lo; t += 2) {
          step(t);
          step(t + 1);
        }
        for (; t + 3 < t_hi; t += 4) {
          const int32_t m0 = steady(t), m1 = steady(t + 1);
          const int32_t m2 = steady(t + 2), m3 = steady(t + 3);
          rmax = max(max(rmax, max(m0, m1)), max(m2, m3));
        }
        for (; t < t_hi; ++t) rmax = max(rmax, steady(t));
      }
      for (; t + 1 < T; t += 2) {
        step(t);
        step(t + 1);
      }
      if (t < T) step(t);
      __builtin_amdgcn_wave_barrier();
      const bool stop_row = real && (float)(rmax + (la - i)) / (float)la < threshold;
      const uint64_t bal = __ballot(stop_row);
      if (bal & seg_lanes) {
        stopped = true;
        break;
      }
    }
    int32_t fh = NW_NEG, fm = 0;
    if (!stopped && lb - la <= bw) {
      const int32_t hn = rb[lb - la + bw].x;
      fh = (hn >> 16) - GO;
      fm = hn & NWK_MMASK;
    }
    if (l == 0) {
      if (stopped) nw_write(p, 0.0f, NW_NEG, threshold, out_sims, out_scores, out_mask);
      else nw_write(p, (float)fm / (float)la, fh, threshold, out_sims, out_scores, out_mask);
    }
  }
}

// ---------------------------------------------------------------------------
// Stripe-pipelined packed-key form.  Every lane walks its stripes back to
// back: stripe g of a segment's pair is group g of L >= max(2S + 1, W) steps
// (L uniform over the wavefront); lane l works on group g from group-local
// step 2l to 2l + L - 1 (positions 0 .. L-1; positions >= W lie outside the
// band and are thrown away), so in the first 2S steps of group g (the
// transition) lanes l > tau/2 still finish group g - 1.  At its switch step
// 2l a lane checks the early stop of its finished row -- an integer bound:
// stop when max matches < x* - la + i, x* the smallest x with fl32(x / la)
// >= threshold -- clears its row maximum and takes its new row.  The last
// row of a group reaches lane 0 of the next group through the row buffer,
// written by lane S-1 (by the lane of row la in the pair's last group, whose
// row-buffer entry lb - la + bw is the result).  Lane 0 reads entry P of it
// at group-local step P - 2 (one step ahead); lane S-1 of the previous group
// wrote it at step P + 2S - 2 - L < P - 2 of this group (L >= 2S + 1), and
// this group's writer overwrites it at step P + 2(writer lane) > P - 2.
// Codes of b are staged per group into one of two buffers (the next group's
// while this one runs).  Transitions whose old and new rows all lie inside
// the band's matrix region (i > bw, i <= lb - bw) run a reduced rule set:
// no column-0 cell, and for a lane still on its old row one band-end test
// (tau <= W - 1 - L + 2l) for the row maximum, the write and (equal) the
// band edge.  After the transition, steps in [t_lo, t_hi) -- every real
// row's lane inside its row's cell range and short of the band's last
// column -- run without rules, unrolled by four.
// Keys carry a diagonal offset: the values of cell (i, j) hold (i + j) << 16
// on top of their scores, and the stored HN, EN and FN sit in the domain of
// the cells that read them (i + j + 1).  A gap extension (-1) is then the
// move to the next domain itself: EN = E and FN = F need no subtraction.
// The diagonal addend carries + 1 << 16 (two domains up from HN's), row 0
// and the column-0 boundary carry their cells' offsets, and the result
// drops la + lb + 1.  Cells outside the matrix (rows past la, columns
// outside [0, lb]) may wrap around; none of them feeds a real cell.
// ---------------------------------------------------------------------------

namespace {

// LDS bytes of one segment of the pipelined kernel (cap >= max(W, 2S + 1)):
// row buffer of cap + 1 int2 (HN, EN), then two staging buffers of
// cap + 2S - 1 code bytes each (16-byte rounded).
__host__ __device__ __forceinline__ int32_t nwp_sb_bytes(int32_t cap, int32_t S) {
  return (cap + 2 * S + 15) & ~15;
}
__host__ __device__ __forceinline__ int32_t nwp_seg_bytes(int32_t cap, int32_t S) {
  return (cap + 1) * 8 + 2 * nwp_sb_bytes(cap, S);
}
__host__ __device__ __forceinline__ int32_t nwp_wave_bytes(int32_t cap16, int32_t cap8) {
  const int32_t a = 4 * nwp_seg_bytes(cap16, 16), b = 8 * nwp_seg_bytes(cap8, 8);
  return a > b ? a : b;
}

__device__ __forceinline__ int32_t nw_wave_max(int32_t v) {
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) v = max(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ int32_t nw_wave_min(int32_t v) {
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) v = min(v, __shfl_xor(v, off));
  return v;
}

// One lane's row of a group.
struct NwRow {
  int32_t xa;      // row residue code (0xFE for rows past la / groups past the pair)
  int32_t tbase;   // its table row (scored form)
  int32_t kr;      // position of column 0 (-1 past row bw)
  int32_t hn_r;    // HN of column 0 (the row's left boundary)
  int32_t klo;     // cell range [klo, klo + kspan] of the row (row maximum)
  uint32_t kspan;
  int32_t need;    // stop when the row maximum is below this (INT_MIN: never)
  bool writer;     // this lane writes its row into the row buffer
  bool real;
  bool inner;      // no column-0 cell and no column past lb in the row (or not real)
};

}  // namespace

template <bool SCORED, int S>
__device__ __forceinline__ void nw_pipe_pairs(
    const int32_t* __restrict__ queue, int32_t qb, int32_t n, int32_t* __restrict__ head,
    const int32_t* __restrict__ pairs, const uint8_t* __restrict__ flat,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ lengths,
    int32_t band_width, float threshold, int32_t raw, const int32_t* __restrict__ tab,
    char* wave_lds, int32_t cap, float* __restrict__ out_sims,
    int32_t* __restrict__ out_scores, uint8_t* __restrict__ out_mask) {
  static_assert(S == 8 || S == 16, "8 or 16 lanes per pair");
  constexpr int PER_WAVE = WAVE / S;
  constexpr int32_t GO = SCORED ? -11 : -10;
  constexpr int32_t GE = -1;
  constexpr int DPP = S == 16 ? 0x111 : 0x112;   // row_shr:1 / row_shr:2
  constexpr int32_t ADD_MATCH = nwk(2 - GO + 1, NWK_PD, 1);
  constexpr int32_t ADD_MISS = nwk(-1 - GO + 1, NWK_PD, 0);
  constexpr int32_t HN_ADD = (GO + 1) * NWK_ONE;
  constexpr int32_t EK_EDGE = nwk(NWK_NEG, NWK_PE, 0);
  constexpr int32_t FN_RESET = nwk(NWK_NEG - 1, NWK_PF, 0);
  constexpr int32_t JUNK_H = nwk(NWK_NEG - 1, 0, 0);
  constexpr int32_t JUNK_E = nwk(NWK_NEG - 1, NWK_PE, 0);
  constexpr int32_t BIG = 1 << 30;
  const int lane = threadIdx.x & (WAVE - 1);
  const int l = S == 16 ? (lane & 15) : ((lane & 15) >> 1);
  const int seg = S == 16 ? (lane >> 4) : ((lane >> 4) * 2 + (lane & 1));
  const uint64_t seg_lanes = S == 16 ? (0xFFFFull << (lane & ~15))
                                     : (0x5555ull << ((lane & ~15) | (lane & 1)));
  const int32_t sbb = nwp_sb_bytes(cap, S);
  char* seg_base = wave_lds + (size_t)seg * nwp_seg_bytes(cap, S);
  int2* rb = reinterpret_cast<int2*>(seg_base);
  uint8_t* sbuf0 = reinterpret_cast<uint8_t*>(seg_base + (size_t)(cap + 1) * sizeof(int2));
  uint8_t* sbuf1 = sbuf0 + sbb;
  for (;;) {
    const int32_t q0 = qb + wave_fetch(head, PER_WAVE);
    if (q0 >= n) break;
    const int32_t slot = q0 + seg;
    const bool valid = slot < n;
    const int64_t p = valid ? queue[slot] : 0;
    NwGeom g;
    g.a_off = 0; g.b_off = 0; g.la = 0; g.lb = 0; g.bw = 0; g.W = 0; g.status = 0;
    if (valid) g = nw_geom(pairs, p, offsets, lengths, band_width, threshold, raw);
    const uint8_t* A = flat + g.a_off;
    const uint8_t* B = flat + g.b_off;
    const int32_t la = g.la, lb = g.lb, bw = g.bw, W = valid ? g.W : 0;
    const int32_t nstr = valid ? (la + S - 1) / S : 0;
    const int32_t L = nw_wave_max(max(2 * S + 1, W));
    const int32_t smax = nw_wave_max(nstr);
    // x*: the smallest x with fl32(x / la) >= threshold, searched around
    // threshold * la (never: 0; above 2: every row stops)
    int32_t xs = 0;
    if (valid && threshold > 0.0f) {
      if (threshold > 2.0f) {
        xs = 2 * la + 2;
      } else {
        const int32_t fx = (int32_t)floorf(threshold * (float)la);
        xs = fx + 3;
        for (int32_t x = max(0, fx - 2); x <= fx + 2; ++x)
          if ((float)x / (float)la >= threshold) { xs = x; break; }
      }
    }
    const int32_t wl = la - 1 - (nstr - 1) * S;
    auto rowp = [&](int32_t gg) {
      NwRow r;
      const bool live = gg >= 0 && gg < nstr;
      const int32_t i = gg * S + l + 1;
      r.real = live && i <= la;
      r.xa = r.real ? (int32_t)A[i - 1] : 0xFE;
      r.tbase = SCORED ? (r.xa < 20 ? r.xa : 20) * NW_TAB : 0;
      r.kr = i <= bw ? bw - i : -1;
      r.hn_r = nwk((i <= bw + 1 ? GO + GE * (i - 1) + i + 1 : NWK_NEG) + GO, 0, 0);
      r.klo = max(0, bw + 1 - i);
      r.kspan = (uint32_t)(min(W - 1, lb - i + bw) - r.klo);
      r.need = r.real ? xs - la + i : INT32_MIN;
      r.writer = live && l == (gg == nstr - 1 ? wl : S - 1);
      r.inner = !r.real || (i > bw && i <= lb - bw);
      return r;
    };
    auto stage = [&](int32_t gg, uint8_t* buf) {
      const bool live = gg < nstr;
      const int32_t jb = gg * S + 1 - bw - (S - 1);
      for (int32_t x = l; x < L + 2 * S - 1; x += S) {
        const int32_t j = jb + x;
        buf[x] = (live && j >= 1 && j <= lb) ? B[j - 1] : (uint8_t)0xFF;
      }
    };
    for (int32_t k = l; k <= L; k += S) rb[k] = nwp_row0(k, bw, lb, GO);
    stage(0, sbuf0);
    for (int32_t x = l; x < L + 2 * S - 1; x += S) sbuf1[x] = 0xFF;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    int32_t hu = JUNK_H, hout = JUNK_H, eout = JUNK_E, fn = FN_RESET, rmax = 0;
    bool stop = false;
    NwRow pr = rowp(-1);
    const int32_t c_l = W - 1 - L + 2 * l;   // old-row band end in transition steps
    for (int32_t gi = 0; gi <= smax; ++gi) {
      const NwRow cr = rowp(gi);
      const uint8_t* sc = (gi & 1) ? sbuf1 : sbuf0;
      const uint8_t* sp = (gi & 1) ? sbuf0 : sbuf1;
      if (l == 0) hu = rb[0].x;
      int2 rn = rb[1];
      const bool fast_ok = pr.inner && cr.inner && (W >= 2 * S + 1 || (!pr.real && !cr.real));
      const bool fast = __ballot(!fast_ok) == 0;

      // One transition step.  FAST: the reduced rule set (see above).
      auto xstep = [&](int32_t tau, auto fast_c) {
        constexpr bool FAST = decltype(fast_c)::value;
        const bool in_c = 2 * l <= tau;
        const bool sw = 2 * l == tau;
        const int2 r = rn;
        rn = rb[min(tau + 2, L)];
        const uint8_t* src = in_c ? sc + (tau - l + S - 1) : sp + (L + tau - l + S - 1);
        const uint32_t v = *src;
        if (sw) {
          stop = stop || rmax < pr.need;
          rmax = 0;
        }
        const int32_t hd = hu;
        hu = __builtin_amdgcn_update_dpp(r.x, hout, DPP, 0xF, 0xF, false);
        const int32_t eu = __builtin_amdgcn_update_dpp(r.y, eout, DPP, 0xF, 0xF, false);
        const int32_t xa = in_c ? cr.xa : pr.xa;
        const int32_t mt = v == (uint32_t)xa ? 1 : 0;
        const int32_t add = SCORED ? tab[(in_c ? cr.tbase : pr.tbase) + (int32_t)(v < 20u ? v : 20u)] + mt
                                   : (mt ? ADD_MATCH : ADD_MISS);
        const int32_t d = hd + add;
        const int32_t k = tau - 2 * l + (in_c ? 0 : L);
        int32_t ek = max(eu, hu) | NWK_PE;
        bool range, wr;
        if constexpr (FAST) {
          const bool valid_old = tau <= c_l;
          if (!in_c && tau == c_l) ek = EK_EDGE;
          range = in_c || valid_old;
          wr = (in_c ? cr.writer : pr.writer) && range;
        } else {
          if (k == W - 1) ek = EK_EDGE;
          range = (uint32_t)(k - (in_c ? cr.klo : pr.klo)) <= (in_c ? cr.kspan : pr.kspan);
          wr = (in_c ? cr.writer : pr.writer) && (uint32_t)k < (uint32_t)W;
        }
        // first cell of a row: the left neighbour is the row's boundary
        const int32_t fk = max(sw ? FN_RESET : fn, sw ? cr.hn_r : hout) | NWK_PF;
        const int32_t h = max(max(d, ek), fk);
        hout = (h & ~NWK_PMASK) + HN_ADD;
        eout = ek;
        fn = fk;
        if constexpr (!FAST) {
          if (k == (in_c ? cr.kr : pr.kr)) {
            hout = in_c ? cr.hn_r : pr.hn_r;
            fn = FN_RESET;
          }
        }
        if (range) rmax = max(rmax, h & NWK_MMASK);
        if (wr) rb[k] = make_int2(hout, eout);
      };
      if (fast) {
        // The reduced rules, fully unrolled: each lane's code pointer, row
        // residue, table row, store base, writer flag and band-end limit
        // switch from its old row to its new one at tau = 2l (one select
        // each on even steps) instead of being chosen per step.  vlim: the
        // old row's cells lie in the band for tau <= c_l (the band edge at
        // equality); the new row's cells always do.
        const uint8_t* cp = sp + (L + S - 1 - l);
        const uint8_t* const cpn = sc + (S - 1 - l);
        int32_t xa = pr.xa;
        int32_t tb = SCORED ? pr.tbase : 0;
        // row-buffer entry of step tau: wp[tau] (old row: k = tau - 2l + L,
        // new row: k = tau - 2l); stores while tau <= wlim (the band-end
        // limit for a writer, -1 otherwise)
        int2* wp = rb + (L - 2 * l);
        int2* const wpn = rb + L - 2 * l - L;
        int32_t wlim = pr.writer ? c_l : -1;
        const int32_t wlim_n = cr.writer ? 0x7FFFFFFF : -1;
        int32_t vlim = c_l;
        const int32_t fks = max(FN_RESET, cr.hn_r) | NWK_PF;   // F of a new row's first cell
#pragma unroll
        for (int tau = 0; tau < 2 * S; ++tau) {
          bool sw = false;
          if ((tau & 1) == 0) {
            sw = l == (tau >> 1);
            stop = stop || (sw && rmax < pr.need);
            rmax = sw ? 0 : rmax;
            cp = sw ? cpn : cp;
            xa = sw ? cr.xa : xa;
            if (SCORED) tb = sw ? cr.tbase : tb;
            wp = sw ? wpn : wp;
            wlim = sw ? wlim_n : wlim;
            vlim = sw ? 0x7FFFFFFF : vlim;
          }
          const int2 r = rn;
          rn = rb[tau + 2];                // tau + 2 <= 2S + 1 <= L
          const uint32_t v = cp[tau];
          const int32_t hd = hu;
          hu = __builtin_amdgcn_update_dpp(r.x, hout, DPP, 0xF, 0xF, false);
          const int32_t eu = __builtin_amdgcn_update_dpp(r.y, eout, DPP, 0xF, 0xF, false);
          const int32_t mt = v == (uint32_t)xa ? 1 : 0;
          const int32_t add = SCORED ? tab[tb + (int32_t)(v < 20u ? v : 20u)] + mt
                                     : (mt ? ADD_MATCH : ADD_MISS);
          const int32_t d = hd + add;
          int32_t ek = max(eu, hu) | NWK_PE;
          if (tau == vlim) ek = EK_EDGE;
          const int32_t fk = sw ? fks : (max(fn, hout) | NWK_PF);
          const int32_t h = max(max(d, ek), fk);
          hout = (h & ~NWK_PMASK) + HN_ADD;
          eout = ek;
          fn = fk;
          if (tau <= vlim) rmax = max(rmax, h & NWK_MMASK);
          if (tau <= wlim) wp[tau] = make_int2(hout, eout);
        }
      } else {
#pragma unroll 2
        for (int32_t tau = 0; tau < 2 * S; ++tau) xstep(tau, std::false_type{});
      }
      // every lane has checked its previous row
      const bool seg_done = ((__ballot(stop) & seg_lanes) != 0) || gi >= nstr;
      if (gi == smax || __ballot(!seg_done) == 0) break;
      stage(gi + 1, (gi & 1) ? sbuf0 : sbuf1);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

      // Body: every lane on its row of this group.
      const int32_t t_lo = max(2 * S, nw_wave_max(cr.real ? cr.klo + 2 * l : 0));
      const int32_t t_hi = min(
          L, nw_wave_min(cr.real ? min(W - 1, cr.klo + (int32_t)cr.kspan + 2 * l + 1) : BIG));
      uint32_t vn = sc[2 * S - l + S - 1];
      auto gstep = [&](int32_t tau) {
        const int32_t k = tau - 2 * l;
        const int2 r = rn;
        const uint32_t v = vn;
        rn = rb[min(tau + 2, L)];
        vn = sc[tau + S - l];
        const int32_t hd = hu;
        hu = __builtin_amdgcn_update_dpp(r.x, hout, DPP, 0xF, 0xF, false);
        const int32_t eu = __builtin_amdgcn_update_dpp(r.y, eout, DPP, 0xF, 0xF, false);
        const int32_t mt = v == (uint32_t)cr.xa ? 1 : 0;
        const int32_t add = SCORED ? tab[cr.tbase + (int32_t)(v < 20u ? v : 20u)] + mt
                                   : (mt ? ADD_MATCH : ADD_MISS);
        const int32_t d = hd + add;
        int32_t ek = max(eu, hu) | NWK_PE;
        if (k == W - 1) ek = EK_EDGE;
        const int32_t fk = max(fn, hout) | NWK_PF;
        const int32_t h = max(max(d, ek), fk);
        hout = (h & ~NWK_PMASK) + HN_ADD;
        eout = ek;
        fn = fk;
        if (k == cr.kr) {
          hout = cr.hn_r;
          fn = FN_RESET;
        }
        if ((uint32_t)(k - cr.klo) <= cr.kspan) rmax = max(rmax, h & NWK_MMASK);
        if (cr.writer && (uint32_t)k < (uint32_t)W) rb[k] = make_int2(hout, eout);
      };
      auto steady = [&](int32_t tau) -> int32_t {
        const int32_t k = tau - 2 * l;
        const int2 r = rn;
        const uint32_t v = vn;
        rn = rb[tau + 2];
        vn = sc[tau + S - l];
        const int32_t hd = hu;
        hu = __builtin_amdgcn_update_dpp(r.x, hout, DPP, 0xF, 0xF, false);
        const int32_t eu = __builtin_amdgcn_update_dpp(r.y, eout, DPP, 0xF, 0xF, false);
        const int32_t mt = v == (uint32_t)cr.xa ? 1 : 0;
        const int32_t add = SCORED ? tab[cr.tbase + (int32_t)(v < 20u ? v : 20u)] + mt
                                   : (mt ? ADD_MATCH : ADD_MISS);
        const int32_t d = hd + add;
        const int32_t ek = max(eu, hu) | NWK_PE;
        const int32_t fk = max(fn, hout) | NWK_PF;
        const int32_t h = max(max(d, ek), fk);
        hout = (h & ~NWK_PMASK) + HN_ADD;
        eout = ek;
        fn = fk;
        if (cr.writer) rb[k] = make_int2(hout, eout);
        return h & NWK_MMASK;
      };
      int32_t tau = 2 * S;
#pragma unroll 1
      for (; tau < t_lo && tau < L; ++tau) gstep(tau);
      for (; tau + 3 < t_hi; tau += 4) {
        const int32_t m0 = steady(tau), m1 = steady(tau + 1);
        const int32_t m2 = steady(tau + 2), m3 = steady(tau + 3);
        rmax = max(max(rmax, max(m0, m1)), max(m2, m3));
      }
      for (; tau < t_hi; ++tau) rmax = max(rmax, steady(tau));
#pragma unroll 1
      for (; tau < L; ++tau) gstep(tau);
      pr = cr;
    }
    const bool stopped = (__ballot(stop) & seg_lanes) != 0;
    if (valid && l == 0) {
      int32_t fh = NW_NEG, fm = 0;
      if (!stopped && lb - la <= bw) {
        const int32_t hn = rb[lb - la + bw].x;
        fh = (hn >> 16) - GO - (la + lb + 1);
        fm = hn & NWK_MMASK;
      }
      if (stopped) nw_write(p, 0.0f, NW_NEG, threshold, out_sims, out_scores, out_mask);
      else nw_write(p, (float)fm / (float)la, fh, threshold, out_sims, out_scores, out_mask);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Both packed-key classes in one persistent launch: every wavefront first
// takes pairs of the narrow class (windows up to w_cap8 positions, 8 lanes
// per pair, queue slots [*e1, *q_end) through *head8), then of the wide class
// (16 lanes, slots [*e0, *e1) through *head16).  The longest narrow pairs,
// handed out first, then overlap the wide class instead of forming a tail of
// their own.  LDS: the table, then one region per wavefront that holds the
// segment buffers of either class.
// PIPE: the stripe-pipelined schedule (nw_pipe_pairs, window capacities
// w_cap16 >= 33 and w_cap8 >= 17) instead of one stripe at a time.
template <bool SCORED, bool PIPE>
__global__ __launch_bounds__(WAVE_BLOCK, PIPE ? 4 : 1) void nw_pack_kernel(
    const int32_t* __restrict__ queue, const int32_t* __restrict__ e0,
    const int32_t* __restrict__ e1, const int32_t* __restrict__ q_end,
    int32_t* __restrict__ head16, int32_t* __restrict__ head8,
    const int32_t* __restrict__ pairs, const uint8_t* __restrict__ flat,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ lengths,
    int32_t band_width, float threshold, int32_t raw, const int8_t* __restrict__ sub,
    int32_t w_cap16, int32_t w_cap8, float* __restrict__ out_sims,
    int32_t* __restrict__ out_scores, uint8_t* __restrict__ out_mask) {
  constexpr int32_t GO = SCORED ? -11 : -10;
  extern __shared__ int4 nw_lds[];
  int32_t* tab = reinterpret_cast<int32_t*>(nw_lds);
  if (SCORED) {
    for (int x = threadIdx.x; x < NW_TAB * NW_TAB; x += blockDim.x) {
      const int r = x / NW_TAB, c = x % NW_TAB;
      const int32_t sc = (r < 20 && c < 20) ? (int32_t)sub[r * 20 + c] : -4;
      tab[x] = nwk(sc - GO + (PIPE ? 1 : 0), NWK_PD, 0);   // pipelined: diagonal offset
    }
  }
  __syncthreads();
  char* wave_lds = reinterpret_cast<char*>(nw_lds) + NW_TAB_WORDS * 4 +
                   (size_t)(threadIdx.x >> 6) * (PIPE ? nwp_wave_bytes(w_cap16, w_cap8)
                                                      : nwk_wave_bytes(w_cap16, w_cap8));
  const int32_t b0 = *e0, b1 = *e1, b2 = *q_end;
  if constexpr (PIPE) {
    nw_pipe_pairs<SCORED, 8>(queue, b1, b2, head8, pairs, flat, offsets, lengths, band_width,
                             threshold, raw, tab, wave_lds, w_cap8, out_sims, out_scores,
                             out_mask);
    nw_pipe_pairs<SCORED, 16>(queue, b0, b1, head16, pairs, flat, offsets, lengths, band_width,
                              threshold, raw, tab, wave_lds, w_cap16, out_sims, out_scores,
                              out_mask);
  } else {
    nw_pack_pairs<SCORED, 8>(queue, b1, b2, head8, pairs, flat, offsets, lengths, band_width,
                             threshold, raw, tab, wave_lds, w_cap8, out_sims, out_scores,
                             out_mask);
    nw_pack_pairs<SCORED, 16>(queue, b0, b1, head16, pairs, flat, offsets, lengths, band_width,
                              threshold, raw, tab, wave_lds, w_cap16, out_sims, out_scores,
                              out_mask);
  }
}

// One pair per thread, one band row of W + 1 int4 per thread in global
// memory (element k of thread x at ws[k * nthreads + x]), updated in place:
// cell k reads the previous row's k (diagonal) and k + 1 (above) before
// overwriting k.  Position W keeps H = E = -inf; the column-0 entry of each
// row is written after the row for the next row's diagonal.
template <bool SCORED>
__global__ __launch_bounds__(256) void nw_plain_kernel(
    const int32_t* __restrict__ list, const int32_t* __restrict__ list_len,
    const int32_t* __restrict__ pairs, const uint8_t* __restrict__ flat,
    const int64_t* __restrict__ offsets, const int32_t* __restrict__ lengths,
    int32_t band_width, float threshold, int32_t raw, const int8_t* __restrict__ sub,
    int4* __restrict__ ws, int32_t ws_w, float* __restrict__ out_sims, int32_t* __restrict__ out_scores,
    uint8_t* __restrict__ out_mask, int32_t* __restrict__ err) {
  constexpr int32_t GO = SCORED ? -11 : -10;
  constexpr int32_t GE = -1;
  const int64_t nth = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t n = *list_len;
  for (int64_t q = tid; q < n; q += nth) {
    const int64_t p = list[q];
    const NwGeom g = nw_geom(pairs, p, offsets, lengths, band_width, threshold, raw);
    const uint8_t* A = flat + g.a_off;
    const uint8_t* B = flat + g.b_off;
    const int32_t la = g.la, lb = g.lb, bw = g.bw, W = g.W;
    if (W + 1 > ws_w) {          // max_len given by the caller was too small
      nw_write(p, 0.0f, NW_NEG, 2.0f, out_sims, out_scores, out_mask);
      atomicOr(err, 1);
      continue;
    }
#define R(k) ws[(int64_t)(k) * nth + tid]
    for (int32_t k = 0; k <= W; ++k) R(k) = nw_row0(k, bw, lb, GO);
    bool stopped = false;
    for (int32_t i = 1; i <= la && !stopped; ++i) {
      const uint32_t xa = A[i - 1];
      const int32_t klo = bw + 1 - i > 0 ? bw + 1 - i : 0;
      const int32_t khi = lb - i + bw < W - 1 ? lb - i + bw : W - 1;
      int32_t hl = (klo == 0 && i == bw + 1) ? GO + GE * (i - 1) : NW_NEG;
      if (klo > 0) hl = GO + GE * (i - 1);   // column 0 lies left of the first cell
      int32_t ml = 0, F = NW_NEG, FM = 0, rmax = 0;
      int4 diag = R(klo);
      for (int32_t k = klo; k <= khi; ++k) {
        const int4 up = R(k + 1);
        const int32_t j = i - bw + k;
        const uint32_t y = B[j - 1];
        const int32_t mt = y == xa ? 1 : 0;
        int32_t sc;
        if (SCORED) sc = (xa < 20u && y < 20u) ? (int32_t)sub[xa * 20 + y] : -4;
        else sc = mt ? 2 : -1;
        const int32_t d = diag.x + sc, dm = diag.y + mt;
        const int32_t eext = up.z + GE, eop = up.x + GO;
        const int32_t e = eext >= eop ? eext : eop;
        const int32_t em = eext >= eop ? up.w : up.y;
        const int32_t fext = k > klo ? F + GE : NW_NEG, fop = hl + GO;
        if (fext >= fop) {
          F = fext;
        } else {
          F = fop;
          FM = ml;
        }
        int32_t h = d, m = dm;
        if (e > h) { h = e; m = em; }
        if (F > h) { h = F; m = FM; }
        R(k) = make_int4(h, m, e, em);
        hl = h;
        ml = m;
        diag = up;
        rmax = m > rmax ? m : rmax;
      }
      if (klo > 0) R(klo - 1) = make_int4(GO + GE * (i - 1), 0, NW_NEG, 0);
      if ((float)(rmax + (la - i)) / (float)la < threshold) stopped = true;
    }
    if (stopped) {
      nw_write(p, 0.0f, NW_NEG, threshold, out_sims, out_scores, out_mask);
    } else if (lb - la > bw) {   // (la, lb) outside the band (raw form only)
      nw_write(p, 0.0f, NW_NEG, threshold, out_sims, out_scores, out_mask);
    } else {
      const int4 f = R(lb - la + bw);
      nw_write(p, (float)f.y / (float)la, f.x, threshold, out_sims, out_scores, out_mask);
    }
#undef R
  }
}

#define NW_WAVE_INST(B_, S_)                                                                   \
  template __global__ void nw_wave_kernel<B_, S_>(                                           \
      const int32_t*, const int32_t*, const int32_t*, int32_t*, const int32_t*, const uint8_t*, \
      const int64_t*, const int32_t*, int32_t, float, int32_t, const int8_t*, int32_t, float*,  \
      int32_t*, uint8_t*);
NW_WAVE_INST(false, 16)
NW_WAVE_INST(true, 16)
#define NW_PACK_INST(B_, P_)                                                                   \
  template __global__ void nw_pack_kernel<B_, P_>(                                           \
      const int32_t*, const int32_t*, const int32_t*, const int32_t*, int32_t*, int32_t*,      \
      const int32_t*, const uint8_t*, const int64_t*, const int32_t*, int32_t, float, int32_t,  \
      const int8_t*, int32_t, int32_t, float*, int32_t*, uint8_t*);
NW_PACK_INST(false, false)
NW_PACK_INST(true, false)
NW_PACK_INST(false, true)
NW_PACK_INST(true, true)
template __global__ void nw_plain_kernel<false>(const int32_t*, const int32_t*, const int32_t*,
    const uint8_t*, const int64_t*, const int32_t*, int32_t, float, int32_t, const int8_t*,
    int4*, int32_t, float*, int32_t*, uint8_t*, int32_t*);
template __global__ void nw_plain_kernel<true>(const int32_t*, const int32_t*, const int32_t*,
    const uint8_t*, const int64_t*, const int32_t*, int32_t, float, int32_t, const int8_t*,
    int4*, int32_t, float*, int32_t*, uint8_t*, int32_t*);

int32_t nw_lds_bytes(int32_t w_cap, int32_t lanes) {
  return NW_TAB_WORDS * 4 + (WAVE_BLOCK / lanes) * nw_seg_bytes(w_cap);
}

int32_t nw_pack_lds_bytes(int32_t w_cap16, int32_t w_cap8, bool pipe) {
  return NW_TAB_WORDS * 4 + (WAVE_BLOCK / WAVE) * (pipe ? nwp_wave_bytes(w_cap16, w_cap8)
                                                       : nwk_wave_bytes(w_cap16, w_cap8));
}

}  // namespace ck
