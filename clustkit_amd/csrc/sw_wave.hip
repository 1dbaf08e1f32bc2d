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
// positions k (column j = i + hint - bw + k), lane l running two steps
// behind lane l-1: at step t lane l is at relative position t - 2l.  The
// cell above (i-1, j) was produced by lane l-1 one step earlier and the
// diagonal (i-1, j-1) two steps earlier, so both arrive through one DPP
// row_shr:1 per step; the cell to the left is the lane's own previous step.
// Lane 0 takes its "above" values from an LDS row buffer (indexed by
// absolute band position) written by lane 15 of the previous stripe.
//
// Position window.  Band positions where every lane of the stripe is left
// of column 1 hold H = 0 and are skipped (the row buffer keeps its "outside"
// values there, which yield the same E); positions where every lane is right
// of the last column feed nothing and are skipped as well.  A stripe then
// runs over [pos0, pos0 + npos) only, which trims the corners of the band
// where it leaves the matrix.
//
// Steps.  When every live window holds at least 34 positions, steps 0..31
// ("ramp up") reset each lane's state at its first cell (lane l at step 2l)
// and mask the running maximum; steps 32 .. (shortest live window, less one
// where it ends at the band edge) have every lane of every live segment past
// its first cell, inside its window and off the band's last column: they
// run without masks, unrolled, with their LDS lookups issued ahead
// ("steady").  The remaining steps ("ramp down") mask the maximum and apply
// the band-edge rule as lanes finish one after another.  Shorter windows run
// every step with every rule.  Lanes outside their window compute
// throw-away values that never enter the maximum or the row buffer.
//
// Arithmetic.  Each DP value is one int32 key (score<<16 | pref<<13 |
// matches, see ck_internal.h) so that a single signed max() both picks the
// winning score and carries the winner's match count with the reference's
// tie preferences.  A steady cell costs ~17 VALU ops (2 DPP moves, 1 add,
// 5 max, or/and/sub normalisations, 1 address, 2 for the running maximum).
// Scores live in the 16-bit half of the key; a pair whose best score
// reaches SAT_LIMIT is handed to the int32 kernel.
//
// Out-of-matrix cells.  Rows past the shorter sequence and columns outside
// 1..lb read the padding code 31, whose substitution score is -16000.  Such
// cells never feed a cell inside the matrix (values flow right and down
// only, and columns < 1 stay at H = 0), and any value they take is bounded
// by an in-matrix value minus a gap penalty, so they can neither raise the
// maximum nor tie it.
#include <type_traits>

#include "ck_internal.h"

namespace ck {

__device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

// A register whose value is left unspecified (no instruction): the lanes
// that skip a masked LDS read keep it, and the row shift that consumes the
// read replaces it in those lanes.
__device__ __forceinline__ int2 unspecified_int2() {
  int2 r;
  asm volatile("" : "=v"(r.x), "=v"(r.y));
  return r;
}

// (a & m) | t in one instruction (the compiler splits it into and + or when
// the mask is a literal).
__device__ __forceinline__ int32_t and_or(int32_t a, int32_t m, int32_t t) {
  int32_t r;
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(m), "s"(t));
  return r;
}

// Row-buffer entries of eight consecutive steps for the first lane of each
// segment (rb[v] at step v), with one LDS read instead of four: lanes 0-3 of
// the segment (of each 8-lane half for S = 8) read two entries each (lane k:
// 2k and 2k+1), and step v's entry reaches the first lane by a shift of v/2
// lanes to the left within the DPP row.  Only the first lane's values are
// defined; the row shift that consumes r2 replaces the others.
__device__ __forceinline__ void rb_block8(const int2* rb, int lr, int2 (&r2)[8]) {
  int2 a = unspecified_int2(), b = unspecified_int2();
  if (lr < 4) {
    a = rb[2 * lr];
    b = rb[2 * lr + 1];
  }
  r2[0] = a;
  r2[1] = b;
#define CK_RB_SHL(K)                                                   \
  r2[2 * K].x = __builtin_amdgcn_mov_dpp(a.x, 0x100 + K, 0xF, 0xF, true);     \
  r2[2 * K].y = __builtin_amdgcn_mov_dpp(a.y, 0x100 + K, 0xF, 0xF, true);     \
  r2[2 * K + 1].x = __builtin_amdgcn_mov_dpp(b.x, 0x100 + K, 0xF, 0xF, true); \
  r2[2 * K + 1].y = __builtin_amdgcn_mov_dpp(b.y, 0x100 + K, 0xF, 0xF, true);
  CK_RB_SHL(1)
  CK_RB_SHL(2)
  CK_RB_SHL(3)
#undef CK_RB_SHL
}

// Row-buffer stores of eight consecutive steps of the last lane of each
// 16-lane segment (w[v] at step v, destined for dst[v]) with one LDS write
// instead of four: lane 15 keeps its values of steps 0 and 1, and those of
// steps 2k, 2k+1 (k = 1..3) move to lane 15 - 4k by DPP row shifts confined
// to that lane's bank of four; lanes 3, 7, 11 and 15 store two entries each.
// (Steps 0 and 1 accumulate the moves: their registers are dead after the
// block, unlike those of step 7, which carry the state into the next one.)
__device__ __forceinline__ void rb_store8(int2* dst, int lr, const int2 (&w)[8]) {
  int2 s0 = w[0], s1 = w[1];
#define CK_RB_GATHER(K, BANK, SHL)                                                  \
  s0.x = __builtin_amdgcn_update_dpp(s0.x, w[2 * K].x, SHL, 0xF, BANK, false);     \
  s0.y = __builtin_amdgcn_update_dpp(s0.y, w[2 * K].y, SHL, 0xF, BANK, false);     \
  s1.x = __builtin_amdgcn_update_dpp(s1.x, w[2 * K + 1].x, SHL, 0xF, BANK, false); \
  s1.y = __builtin_amdgcn_update_dpp(s1.y, w[2 * K + 1].y, SHL, 0xF, BANK, false);
  CK_RB_GATHER(1, 0x4, 0x104)   // lane 11 <- lane 15, row_shl:4: steps 2, 3
  CK_RB_GATHER(2, 0x2, 0x108)   // lane 7  <- lane 15, row_shl:8: steps 4, 5
  CK_RB_GATHER(3, 0x1, 0x10C)   // lane 3  <- lane 15, row_shl:12: steps 6, 7
#undef CK_RB_GATHER
  if ((lr & 3) == 3) {
    int2* q = dst + (6 - 2 * (lr >> 2));
    q[0] = s0;
    q[1] = s1;
  }
}

// Key layouts.  Narrow (the main pass): score << 16 | pref(3 bits) << 13 |
// matches, preferences 0 (gap open) < F < E < diagonal < zero, scores below
// 32700.  Wide (the re-run of pairs that reach 32700): score << 15 |
// pref(2 bits) << 13 | matches, preferences 0 < F < E < diagonal, the zero
// candidate applied by a compare instead of a key, scores below 65500.
template <bool WIDE>
struct KeyLayout {
  static constexpr int SHIFT = WIDE ? 15 : 16;
  static constexpr int32_t ONE = 1 << SHIFT;
  static constexpr int32_t LOW = ONE - 1;
  static constexpr int32_t PMASK = (WIDE ? 3 : 7) << 13;
  static constexpr int32_t SAT = WIDE ? 65500 : SAT_LIMIT;
  static constexpr int32_t key(int32_t score, int32_t pref, int32_t m) {
    return (int32_t)((uint32_t)score << SHIFT) + pref + m;
  }
  static constexpr int32_t HN0 = key(-11, 0, 0);              // H = 0
  static constexpr int32_t ENN = key(KEY_NEG_SCORE, PREF_E, 0);
  static constexpr int32_t FNN = key(KEY_NEG_SCORE, PREF_F, 0);
  static constexpr int32_t ENE = key(-12, PREF_E, 0);         // E beyond the band edge
  static constexpr int32_t K0 = WIDE ? 0 : key(-1, PREF_0, 0);
  static constexpr int32_t ZERO_M1 = key(-1, 0, 0);           // H = 0, minus-one domain
};

// Per-lane DP state of one stripe.
struct LaneState {
  int32_t hu;    // HN of the cell above, received this step
  int32_t hout;  // HN of this lane's last cell (read by lane + 1)
  int32_t eout;  // EN of this lane's last cell
  int32_t hl;    // HN of the cell to the left
  int32_t fl;    // FN of the cell to the left
  int32_t best;  // first maximum of this lane's row (HN form)
};

// Step variants: which per-lane conditions a step has to honour.
//   RESET  a lane's first position (relative 0) starts from "outside" state
//   EDGE   at the band's last column E comes from outside the band
//   MASK   only positions inside the window enter the running maximum
//   WRITE  lane 15 stores its cell into the row buffer
//   NOBEST the running maximum is left to the caller (chunked first maximum)
//   BESTH  the running maximum holds h with its preference cleared (HN + 10)
enum : int { M_RESET = 1, M_EDGE = 2, M_MASK = 4, M_WRITE = 8, M_NOBEST = 16, M_BESTH = 32 };
constexpr int GENERIC = M_RESET | M_EDGE | M_MASK | M_WRITE;
constexpr int RAMP_UP = M_RESET | M_MASK | M_WRITE;
constexpr int STEADY_NOWRITE = 0;
constexpr int RAMP_DOWN = M_EDGE | M_MASK | M_WRITE;

// One DP step.  tv: diagonal addend of this lane's cell; rb: row-buffer
// entry for lane 0 (cell above of the stripe's first row); kp: relative
// position of this lane; npos/edge: window length and the relative position
// of the band's last column (-1 when outside the window); wr: this lane
// stores its cell into *wslot.  ROFF: keys of row i carry + i << 16 (E then
// needs no gap-extension subtraction; the table addend is s + 11 and the
// zero candidate is the lane's k0 = K0 + i << 16).
template <int MODE, int S, bool WIDE, bool ROFF = false>
__device__ __forceinline__ void dp_step(LaneState& s, int32_t tv, int2 rb, int32_t kp,
                                        int32_t npos, int32_t edge, bool wr, int2* wslot,
                                        bool mid, int32_t k0 = KeyLayout<WIDE>::K0) {
  using L = KeyLayout<WIDE>;
  // Row shift by one lane: within 16-lane rows (row_shr:1), or across the
  // whole wavefront for 64-lane chains (wave_shr:1); the first lane keeps rb.
  constexpr int DPP = S == 64 ? 0x138 : 0x111;
  const int32_t hd = s.hu;
  s.hu = __builtin_amdgcn_update_dpp(rb.x, s.hout, DPP, 0xF, 0xF, false);
  int32_t eu = __builtin_amdgcn_update_dpp(rb.y, s.eout, DPP, 0xF, 0xF, false);
  if (S == 8 && mid) {     // first lane of the second segment of a DPP row
    s.hu = rb.x;
    eu = rb.y;
  }
  if ((MODE & M_RESET) && kp == 0) {
    s.hl = L::HN0;
    s.fl = L::FNN;
  }
  const int32_t d = hd + tv;
  int32_t e = imax(eu, s.hu) | PREF_E;
  if (!ROFF) e -= L::ONE;
  if ((MODE & M_EDGE) && kp == edge) e = L::ENE;
  const int32_t f = (imax(s.fl, s.hl) | PREF_F) - L::ONE;
  int32_t h;
  if (WIDE) {
    h = imax(imax(d, e), f);
    if (h < 0) h = L::ZERO_M1;            // best candidate <= 0: H = 0, no matches
  } else {
    h = imax(imax(d, e), imax(f, ROFF ? k0 : L::K0));
  }
  const int32_t hm = h & ~L::PMASK;
  const int32_t hn = hm - 10 * L::ONE;
  if (!(MODE & M_NOBEST)) {
    const int32_t bv = (MODE & M_BESTH) ? hm : hn;
    bool upd = WIDE ? bv > (s.best | L::LOW) : (bv >> L::SHIFT) > (s.best >> L::SHIFT);
    if (MODE & M_MASK) upd = upd && (uint32_t)kp < (uint32_t)npos;
    if (upd) s.best = bv;
  }
  s.hl = hn;
  s.fl = f;
  s.hout = hn;
  s.eout = e;
  if ((MODE & M_WRITE) && wr) *wslot = make_int2(hn, e);
}

// Dynamic LDS layout per block:
//   int32  table[32][33]       diagonal addends (score+10)<<16 | pref | match,
//                              rows padded to 33 words (bank spread)
//   per segment (16 per block):
//     int2  rowbuf[rb_stride]  (HN, EN) of the previous stripe's last row,
//                              by absolute band position
//     uint8 bstage[bs_stride]  4 * residue code of the stripe's columns
//
// Work: pairs queue[0 .. *q_len), handed out through *q_head; pairs this
// pass cannot finish (score at the layout's limit, raw codes > 30) are
// appended to fail_list (*fail_len).
template <int S, bool WIDE>
__global__ __launch_bounds__(WAVE_BLOCK) void sw_wave_kernel(
    const int32_t* __restrict__ queue, int32_t* __restrict__ q_head,
    const int32_t* __restrict__ q_len, const int32_t* __restrict__ pairs,
    const uint8_t* __restrict__ flat, const int64_t* __restrict__ offsets,
    const int32_t* __restrict__ lengths, const int32_t* __restrict__ hints, int32_t band,
    const int8_t* __restrict__ sub, float threshold, uint32_t flags, int32_t rb_stride,
    int32_t bs_stride, float* __restrict__ out_sims, int32_t* __restrict__ out_scores,
    int32_t* __restrict__ out_matches, uint8_t* __restrict__ out_mask,
    int32_t* __restrict__ fail_list, int32_t* __restrict__ fail_len) {
  using L = KeyLayout<WIDE>;
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* table = reinterpret_cast<int32_t*>(smem);
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  const int seg = lane / S;
  const int lr = lane & (S - 1);
  const int sidx = wv * (WAVE / S) + seg;
  int2* rowbuf = reinterpret_cast<int2*>(smem + TABLE_BYTES) + sidx * rb_stride;
  uint8_t* bst = smem + TABLE_BYTES +
                 (WAVE_BLOCK / S) * rb_stride * (int)sizeof(int2) + sidx * bs_stride;
  const bool writer = lr == S - 1;
  const bool mid = S == 8 && (lane & 15) == 8;
  constexpr int32_t RAMP = 2 * (S - 1);              // the last lane starts here
  constexpr int32_t RAMP_UP_STEPS = RAMP + 2;         // whole blocks of 8

  for (int x = threadIdx.x; x < TABLE_DIM * TABLE_DIM; x += blockDim.x) {
    const int a = x >> 5, b = x & 31;
    int32_t s, m = 0;
    if (a == CODE_POISON || b == CODE_POISON) {
      s = POISON_SCORE;
    } else {
      s = (a < 20 && b < 20) ? (int32_t)sub[a * 20 + b] : UNKNOWN_SCORE;
      m = (a == b) ? 1 : 0;
    }
    table[a * TABLE_STRIDE + b] = L::key(s + 10, PREF_D, m);
  }
  __syncthreads();

  const int32_t qlen = *q_len;
  for (;;) {
    const int32_t base = wave_fetch(q_head, WAVE / S);
    if (base >= qlen) break;
    const int32_t q = base + seg;
    const bool valid = q < qlen;
    const int64_t p = valid ? (int64_t)queue[q] : 0;
    PairGeom g;
    g.a_off = 0; g.b_off = 0; g.la = 0; g.lb = 0; g.bw = 0; g.hint = 0;
    if (valid) g = pair_geom(pairs, p, offsets, lengths, hints, band);
    const int32_t W = valid ? 2 * g.bw + 1 : 0;
    const int32_t nstr = valid ? (g.la + S - 1) / S : 0;
    int32_t wmax = W, smax = nstr;
#pragma unroll
    for (int off = S; off < WAVE; off <<= 1) {
      wmax = imax(wmax, __shfl_xor(wmax, off));
      smax = imax(smax, __shfl_xor(smax, off));
    }
    for (int x = lr; x < wmax + RAMP + 1; x += S) rowbuf[x] = make_int2(L::HN0, L::ENN);
    int32_t seg_best = L::HN0;
    int32_t bad = 0;
    const uint8_t* A = flat + g.a_off;
    const uint8_t* B = flat + g.b_off;

    for (int32_t s = 0; s < smax; ++s) {
      const bool live = s < nstr;
      // Window of band positions: skip positions where every lane is left
      // of column 1 or right of column lb.
      const int32_t c0 = s * S + 1 + g.hint - g.bw;        // lane 0's column at k = 0
      const int32_t kend = min(W - 1, g.lb - c0);
      int32_t pos0 = live ? max(0, 1 - (c0 + S - 1)) : 0;
      const int32_t npos = live ? max(0, kend - pos0 + 1) : 0;
      if (npos == 0) pos0 = 0;                                // keep row-buffer reads in range
      const int32_t edge = (npos > 0 && kend == W - 1) ? npos - 1 : -1;
      // tmax: longest window; tmin: shortest live window; wsteady: steps
      // before which lane 0 of every live segment is inside its window and
      // short of the band's last column.
      int32_t tmax = npos, tmin = live ? npos : 0x7FFFFFFF;
      int32_t wsteady = live ? (edge >= 0 ? npos - 1 : npos) : 0x7FFFFFFF;
#pragma unroll
      for (int off = S; off < WAVE; off <<= 1) {
        tmax = imax(tmax, __shfl_xor(tmax, off));
        tmin = min(tmin, __shfl_xor(tmin, off));
        wsteady = min(wsteady, __shfl_xor(wsteady, off));
      }
      const int32_t T = tmax + RAMP;

      // Row residue per lane and 4 * column residue per staged position
      // x = j - (c0 + pos0) + 15.
      const int32_t i = s * S + lr + 1;
      int32_t acode = CODE_POISON;
      if (live && i <= g.la) {
        const int32_t r = A[i - 1];
        bad |= r > CODE_MAX_FAST;
        acode = r > CODE_MAX_FAST ? CODE_MAX_FAST : r;
      }
      const int32_t cb = c0 + pos0 - (S - 1);
      for (int x = lr; x < T + S; x += S) {
        const int32_t j = cb + x;
        int32_t code = CODE_POISON;
        if (live && j >= 1 && j <= g.lb) {
          const int32_t r = B[j - 1];
          bad |= r > CODE_MAX_FAST;
          code = r > CODE_MAX_FAST ? CODE_MAX_FAST : r;
        }
        bst[x] = (uint8_t)(code * 4);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

      const int2* rb = rowbuf + pos0 + 1;              // lane 0 reads rb[t]
      int2* rw = rowbuf + pos0 - RAMP;                 // lane 15 writes rw[t]
      const uint8_t* bl = bst + (S - 1) - lr;        // this lane's code at step t: bl[t]
      const unsigned char* tab = reinterpret_cast<const unsigned char*>(table + acode * TABLE_STRIDE);
      auto tv_at = [&](int32_t t) {
        return *reinterpret_cast<const int32_t*>(tab + bl[t]);
      };
      LaneState st;
      st.hu = rowbuf[pos0].x;
      st.hout = L::HN0;
      st.eout = L::ENN;
      st.hl = L::HN0;
      st.fl = L::FNN;
      st.best = L::HN0;
      const int32_t n_l = npos;
      const int32_t e_l = edge;

      // Lane 0's row-buffer read at step t, kept inside this segment's
      // written range once the window is done (values there are unused).
      auto rb_at = [&](int32_t t) { return rb[min(t, n_l)]; };
      int32_t t = 0;
      if (tmin >= RAMP_UP_STEPS + 2) {
        // Ramp up: lanes start one after another (lane l at step 2l, lane 15
        // at step 30 and 31 inside the ramp, so every lane's first cell takes
        // the reset); every window is longer than the ramp plus the band
        // edge, so no lane reaches its end or the edge here.  Lane 15 stores
        // its first two cells (steps 30, 31) into the row buffer.
        {
          constexpr int U = 8;
#pragma unroll 1
          for (; t < RAMP_UP_STEPS - U; t += U) {
            int32_t tv[U];
            int2 r2[U];
#pragma unroll
            for (int u = 0; u < U; ++u) tv[u] = tv_at(t + u);
#pragma unroll
            for (int u = 0; u < U; ++u) r2[u] = rb[t + u];
#pragma unroll
            for (int u = 0; u < U; ++u)
              dp_step<RAMP_UP & ~M_WRITE, S, WIDE>(st, tv[u], r2[u], t + u - 2 * lr, n_l, e_l, false,
                                          nullptr, mid);
          }
          constexpr int B0 = RAMP_UP_STEPS - U;     // last block: steps 24..31
          int32_t tv[U];
          int2 r2[U];
#pragma unroll
          for (int u = 0; u < U; ++u) tv[u] = tv_at(B0 + u);
#pragma unroll
          for (int u = 0; u < U; ++u) r2[u] = rb[B0 + u];
#pragma unroll
          for (int u = 0; u < U; ++u)
            dp_step<RAMP_UP, S, WIDE>(st, tv[u], r2[u], B0 + u - 2 * lr, n_l, e_l,
                             writer && B0 + u >= RAMP, rw + B0 + u, mid);
          t = RAMP_UP_STEPS;
        }
        // Steady: every lane of every live segment inside its window, past
        // its first cell and short of the band edge.
        {
          constexpr int U = 8;
          for (; t + U <= wsteady; t += U) {
            int32_t tv[U];
            int2 r2[U];
#pragma unroll
            for (int u = 0; u < U; ++u) tv[u] = tv_at(t + u);
#pragma unroll
            for (int u = 0; u < U; ++u) r2[u] = rb[t + u];
            // Lane S-1's cells of the block go to the row buffer together
            // (paired 16-byte stores under one exec mask).
            int2 wv_[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              dp_step<STEADY_NOWRITE, S, WIDE>(st, tv[u], r2[u], 0, n_l, e_l, false, nullptr, mid);
              wv_[u] = make_int2(st.hout, st.eout);
            }
            if (writer) {
#pragma unroll
              for (int u = 0; u < U; ++u) rw[t + u] = wv_[u];
            }
          }
        }
        // Ramp down (and the steady remainder): lanes finish one after another.
#pragma unroll 2
        for (; t < T; ++t) {
          const bool wr = writer && (uint32_t)(t - RAMP) < (uint32_t)n_l;
          dp_step<RAMP_DOWN, S, WIDE>(st, tv_at(t), rb_at(t), t - 2 * lr, n_l, e_l, wr, rw + t, mid);
        }
      } else {
        // Short windows (or a live segment with an empty one): every step
        // honours starts, ends and the band edge.
#pragma unroll 2
        for (; t < T; ++t) {
          const bool wr = writer && (uint32_t)(t - RAMP) < (uint32_t)n_l;
          dp_step<GENERIC, S, WIDE>(st, tv_at(t), rb_at(t), t - 2 * lr, n_l, e_l, wr, rw + t, mid);
        }
      }

      // First maximum of the stripe: highest score, then lowest row (lane).
      int32_t r = (st.best & ~L::LOW) | (S - 1 - lr);
#pragma unroll
      for (int off = S / 2; off >= 1; off >>= 1) r = imax(r, __shfl_xor(r, off));
      const int32_t wl = (S - 1) - (r & (S - 1));
      const int32_t wbest = __shfl(st.best, (lane & ~(S - 1)) | wl);
      if (live && (wbest >> L::SHIFT) > (seg_best >> L::SHIFT)) seg_best = wbest;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    int32_t any_bad = bad;
#pragma unroll
    for (int off = 1; off < S; off <<= 1) any_bad |= __shfl_xor(any_bad, off);
    if (valid && lr == 0) {
      int32_t score = (seg_best >> L::SHIFT) + 11;
      int32_t matches = seg_best & MATCH_MASK;
      if (any_bad || score >= L::SAT) {
        const int32_t slot = atomicAdd(fail_len, 1);
        fail_list[slot] = (int32_t)p;
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

// Stripe-pipelined form (16-lane segments, four pairs per wavefront, narrow
// keys).  Every lane walks its stripes back to back: lane 0 starts stripe
// g+1 right after its last cell of stripe g, while lanes 1-15 finish stripe
// g, so the 2S-2 ramp steps of a stripe overlap the next one instead of
// idling.  Stripe g of the four pairs forms group g, which lasts
// L_g = max(32, longest window of the group) steps for lane 0.
//
// Group g, local step tau (lane l is at window position kp = tau - 2l):
//   tau in [0, 32)   transition: lane l switches from group g-1 to group g
//                    at tau = 2l (row code, staging buffer, window, band
//                    edge, write base; F and H to the left reset; its running
//                    maximum of group g-1 saved for the reduction); lanes not
//                    yet switched finish group g-1 at kp = L_{g-1} + tau - 2l
//   then             the maxima of group g-1 are reduced; group g+1's column
//                    codes are staged into the other buffer
//   tau in [32, wsteady)   steady: every lane of every live window inside
//                    its window and short of the band edge (as in
//                    sw_wave_kernel)
//   tau in [wsteady, L_g)  tail: masks, band edge, guarded row-buffer stores
// A last, empty group drains the final stripe.
//
// Lane 0 reads the row above from its segment's row buffer, written by lane
// 15 one group earlier: position P is read at group-local step P - pos0_g - 1
// of group g and was written at step P - pos0_{g-1} + 30 of group g-1, at
// least L_{g-1} - 31 >= 1 steps earlier (pos0 never grows from one stripe to
// the next); lane 15 of group g overwrites P only 31 steps after lane 0 of
// group g read it.  All steady-block stores of group g-1 are issued before
// group g starts.  Lane 0's diagonal at its first cell of a group is the row
// buffer entry at pos0_g (the other lanes receive theirs from lane l-1,
// which switched two steps earlier).  Per pair and stripe this runs
// max(32, window) steps instead of window + 30.
//
// Row offsets (narrow keys).  Every key of row i carries i << 16 on top of
// its score: E = max(E_up - 1, H_up - 11) of row i from row i-1's keys is then
// max(EN_up, HN_up) with no subtraction (the gap extension is the offset
// step), the diagonal addend becomes (s + 11) << 16, and the zero candidate
// and the reset values of H and F follow the lane's row (k0 = K0 + i << 16,
// switched with the row).  Comparisons within a row are unchanged; the
// per-row maxima drop the offset in the reduction.  Three rules keep the
// offsets consistent: the row buffer's entries left of column 1 (never
// written by the stripe above) are refilled with H = 0 of the row above the
// next stripe before it reads them; a lane's last cell of the previous group
// in a transition (step 2l - 1) takes the band-edge E, because the cell
// above it is computed by lane l-1 already in the new group, S - 1 rows
// further down (that cell is at the band's last column, right of the
// matrix or outside the window, so its E is not needed); and a pair whose
// offset row maximum reaches the limit (score + row >= 32700) is handed to
// the wide pass, so no key of an in-matrix cell can wrap.
//
// Reduced transitions.  When every segment's previous window spans its
// group (npos = L) and every new window is longer than 2S positions, no cell
// of the transition is outside its window and the band-edge rule only falls
// on the cells of step 2l - 1 (where the rule above applies anyway): these
// transitions run without masks, with the code and table pointers switched
// per lane instead of selected per lookup.
// MINB: workgroups per CU the register allocation is sized for (4 where the
// LDS footprint allows four, i.e. window capacities up to 129 positions).
// SHORTW: built for bands whose windows are shorter than 2S + FB - 1
// positions, where reduced transitions read and store the row buffer step
// by step; the other builds carry only the batched form (one more unrolled
// transition costs the long-window kernel ~1% at band 100).
template <int S, bool WIDE, int MINB, int NT, bool SHORTW>
__global__ __launch_bounds__(NT, MINB) void sw_pipe_kernel(
    const int32_t* __restrict__ queue, int32_t* __restrict__ q_head,
    const int32_t* __restrict__ q_len, const int32_t* __restrict__ pairs,
    const uint8_t* __restrict__ flat, const int64_t* __restrict__ offsets,
    const int32_t* __restrict__ lengths, const int32_t* __restrict__ hints, int32_t band,
    const int8_t* __restrict__ sub, float threshold, uint32_t flags, int32_t rb_stride,
    int32_t bs_stride, float* __restrict__ out_sims, int32_t* __restrict__ out_scores,
    int32_t* __restrict__ out_matches, uint8_t* __restrict__ out_mask,
    int32_t* __restrict__ fail_list, int32_t* __restrict__ fail_len) {
  using L = KeyLayout<WIDE>;
  constexpr bool ROFF = !WIDE;           // row offsets on the narrow keys
  // Narrow keys: the lanes' running maxima hold h with its preference
  // cleared, i.e. HN + 10 << 16 (BHO), which the steady blocks tag for free.
  constexpr int32_t BHO = ROFF ? 10 * L::ONE : 0;
  constexpr int32_t RAMP = 2 * (S - 1);
  constexpr int32_t XW = 2 * S;          // transition steps
  constexpr int32_t STAGE_PAD = XW + S;  // staged codes beyond L_g
  constexpr int DPP = S == 64 ? 0x138 : 0x111;
  constexpr int XU = S <= 16 ? XW : 4;   // transition unroll
  constexpr int FB = 8;                   // block of the reduced transition
  extern __shared__ __align__(16) unsigned char smem[];
  int32_t* table = reinterpret_cast<int32_t*>(smem);
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  const int seg = lane / S;
  const int lr = lane & (S - 1);
  const int sidx = wv * (WAVE / S) + seg;
  int2* rowbuf = reinterpret_cast<int2*>(smem + TABLE_BYTES) + sidx * rb_stride;
  // Two staging buffers per segment (groups alternate), as byte offsets.
  const int32_t bs0 = TABLE_BYTES + (NT / S) * rb_stride * (int)sizeof(int2) +
                      sidx * 2 * bs_stride;
  const int32_t bs1 = bs0 + bs_stride;
  const bool writer = lr == S - 1;
  const bool mid = S == 8 && (lane & 15) == 8;

  for (int x = threadIdx.x; x < TABLE_DIM * TABLE_DIM; x += blockDim.x) {
    const int a = x >> 5, b = x & 31;
    int32_t sc, m = 0;
    if (a == CODE_POISON || b == CODE_POISON) {
      sc = POISON_SCORE;
    } else {
      sc = (a < 20 && b < 20) ? (int32_t)sub[a * 20 + b] : UNKNOWN_SCORE;
      m = (a == b) ? 1 : 0;
    }
    table[a * TABLE_STRIDE + b] = L::key(sc + (ROFF ? 11 : 10), PREF_D, m);
  }
  __syncthreads();

  const int32_t qlen = *q_len;
  for (;;) {
    const int32_t base = wave_fetch(q_head, WAVE / S);
    if (base >= qlen) break;
    const int32_t q = base + seg;
    const bool valid = q < qlen;
    const int64_t p = valid ? (int64_t)queue[q] : 0;
    PairGeom g;
    g.a_off = 0; g.b_off = 0; g.la = 0; g.lb = 0; g.bw = 0; g.hint = 0;
    if (valid) g = pair_geom(pairs, p, offsets, lengths, hints, band);
    const int32_t W = valid ? 2 * g.bw + 1 : 0;
    const int32_t nstr = valid ? (g.la + S - 1) / S : 0;
    int32_t wmax = W, smax = nstr;
#pragma unroll
    for (int off = S; off < WAVE; off <<= 1) {
      wmax = imax(wmax, __shfl_xor(wmax, off));
      smax = imax(smax, __shfl_xor(smax, off));
    }
    // Every entry a transition may read (up to pos0 + 2S <= wmax + RAMP + 1)
    // holds this pair's values.
    for (int x = lr; x < wmax + RAMP + 2; x += S) rowbuf[x] = make_int2(L::HN0, L::ENN);
    int32_t seg_best = L::HN0;
    int32_t bad = 0;
    int32_t satf = 0;   // an offset row maximum reached the limit
    const uint8_t* A = flat + g.a_off;
    const uint8_t* B = flat + g.b_off;

    // Window of stripe s of this segment's pair; false when s is past it.
    auto window = [&](int32_t s, int32_t& pos0, int32_t& npos, int32_t& edge) -> bool {
      const bool live = s < nstr;
      const int32_t c0 = s * S + 1 + g.hint - g.bw;
      const int32_t kend = min(W - 1, g.lb - c0);
      pos0 = live ? max(0, 1 - (c0 + S - 1)) : 0;
      npos = live ? max(0, kend - pos0 + 1) : 0;
      if (npos == 0) pos0 = 0;
      edge = (npos > 0 && kend == W - 1) ? npos - 1 : -1;
      return live;
    };
    // 4 * column code of stripe s at x = j - (c0 + pos0) + 15, x < nx, into
    // the buffer at byte offset bo; returns this lane's row code.
    auto stage = [&](int32_t s, bool live, int32_t pos0, int32_t bo, int32_t nx) -> int32_t {
      const int32_t cb = s * S + 1 + g.hint - g.bw + pos0 - (S - 1);
      // Four residues per lane in flight: every load reads a clamped, valid
      // address (flat holds at least one byte), out-of-range columns are
      // replaced by the padding code afterwards.
      const int32_t jhi = max(g.lb, 1);
      constexpr int NL = 4;
      for (int x0 = lr; x0 < nx; x0 += NL * S) {
        int32_t r[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
          const int32_t j = min(max(cb + x0 + k * S, 1), jhi);
          r[k] = B[j - 1];
        }
#pragma unroll
        for (int k = 0; k < NL; ++k) {
          const int32_t x = x0 + k * S;
          const int32_t j = cb + x;
          const bool in = live && j >= 1 && j <= g.lb;
          bad |= in && r[k] > CODE_MAX_FAST;
          const int32_t code = in ? (r[k] > CODE_MAX_FAST ? CODE_MAX_FAST : r[k]) : CODE_POISON;
          if (x < nx) smem[bo + x] = (uint8_t)(code * 4);
        }
      }
      const int32_t i = s * S + lr + 1;
      int32_t acode = CODE_POISON;
      if (live && i <= g.la) {
        const int32_t r = A[i - 1];
        bad |= r > CODE_MAX_FAST;
        acode = r > CODE_MAX_FAST ? CODE_MAX_FAST : r;
      }
      return acode;
    };
    // Stripe s when stripe s-1 was staged at byte offset bo_p (np codes) and
    // neither window starts left of the matrix: the band moves S columns per
    // stripe, so code x of stripe s is code x + S of stripe s-1.  Sixteen
    // bytes per lane are copied within LDS; only the columns past the old
    // buffer's end are loaded (and checked) from the sequence.
    auto stage_shift = [&](int32_t s, int32_t bo, int32_t nx, int32_t bo_p, int32_t np) -> int32_t {
      // chunks of CB bytes (16, or 8 for 8-lane segments: the source is
      // then only 8-byte aligned), whole chunks inside both buffers
      constexpr int CB = S % 16 == 0 ? 16 : 8;
      const int32_t ncopy = min(np - S, nx) & ~(CB - 1);
      for (int32_t x = lr * CB; x < ncopy; x += S * CB) {
        if constexpr (CB == 16)
          *reinterpret_cast<int4*>(smem + bo + x) =
              *reinterpret_cast<const int4*>(smem + bo_p + S + x);
        else
          *reinterpret_cast<int2*>(smem + bo + x) =
              *reinterpret_cast<const int2*>(smem + bo_p + S + x);
      }
      const int32_t cb = s * S + 1 + g.hint - g.bw - (S - 1);
      for (int32_t x = ncopy + lr; x < nx; x += S) {
        const int32_t j = cb + x;
        int32_t code = CODE_POISON;
        if (j >= 1 && j <= g.lb) {
          const int32_t r = B[j - 1];
          bad |= r > CODE_MAX_FAST;
          code = r > CODE_MAX_FAST ? CODE_MAX_FAST : r;
        }
        smem[bo + x] = (uint8_t)(code * 4);
      }
      const int32_t i = s * S + lr + 1;
      int32_t acode = CODE_POISON;
      if (i <= g.la) {
        const int32_t r = A[i - 1];
        bad |= r > CODE_MAX_FAST;
        acode = r > CODE_MAX_FAST ? CODE_MAX_FAST : r;
      }
      return acode;
    };
    // Maximum / minimum over the wavefront of a value that is equal in
    // every lane of a segment: one readlane per segment instead of LDS
    // permutes (every lane is active here).
    auto wave_max = [&](int32_t v) {
      int32_t r = __builtin_amdgcn_readlane(v, 0);
#pragma unroll
      for (int q = S; q < WAVE; q += S) r = imax(r, __builtin_amdgcn_readlane(v, q));
      return r;
    };
    auto wave_min = [&](int32_t v) {
      int32_t r = __builtin_amdgcn_readlane(v, 0);
#pragma unroll
      for (int q = S; q < WAVE; q += S) r = min(r, __builtin_amdgcn_readlane(v, q));
      return r;
    };

    // Group 0 staged; the view every lane starts from is an empty group
    // (poison codes in buffer 1, empty window, previous length 32).
    int32_t pos0_n, npos_n, edge_n;
    bool live_n = window(0, pos0_n, npos_n, edge_n);
    int32_t L_n = wave_max(imax(XW, npos_n));
    int32_t ac_n = stage(0, live_n, pos0_n, bs0, L_n + STAGE_PAD);
    for (int x = lr; x < XW + STAGE_PAD; x += S) smem[bs1 + x] = (uint8_t)(CODE_POISON * 4);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // The previous group's view: at first an empty group of 32 steps over
    // poison codes (buffer 1).  blo: byte offset of this lane's code at
    // group-local step 0 (code of step tau at smem[blo + tau]); tabo: byte
    // offset of this lane's table row.
    int32_t blo_p = bs1 + (S - 1) - lr;
    int32_t tabo_p = CODE_POISON * TABLE_STRIDE * (int32_t)sizeof(int32_t);
    int32_t pos0_p = 0, npos_p = 0, edge_p = -1, L_p = XW;
    LaneState st;
    st.hu = L::HN0;
    st.hout = L::HN0;
    st.eout = L::ENN;
    st.hl = L::HN0;
    st.fl = L::FNN;
    st.best = L::HN0 + BHO;
    int32_t pbest = L::HN0 + BHO;
    bool live_prev = false;
    // F of a lane's first cell in a group (left neighbour outside the band).
    constexpr int32_t F_FIRST = ((L::FNN > L::HN0 ? L::FNN : L::HN0) | PREF_F) - L::ONE;
    // Zero candidate of this lane's row (row offsets): the empty group
    // before group 0 stands for row lr + 1 - S.
    int32_t k0 = ROFF ? L::K0 + (lr + 1 - S) * L::ONE : L::K0;
    constexpr int32_t SAT_KEY = L::key(L::SAT - 11, 0, 0);

    for (int32_t gi = 0; gi <= smax; ++gi) {
      // This group's window (segment-uniform) and this lane's row code.
      const int32_t pos0_c = pos0_n, npos_c = npos_n, edge_c = edge_n, L_c = L_n;
      const bool live_c = live_n;
      const int32_t tabo_c = ac_n * TABLE_STRIDE * (int32_t)sizeof(int32_t);
      const int32_t blo_c = ((gi & 1) ? bs1 : bs0) + (S - 1) - lr;
      const int32_t rbase = pos0_c + 1;     // lane 0 reads rowbuf[rbase + tau]
      const int32_t hdiag = rowbuf[pos0_c].x;
      if (lr == 0) st.hu = hdiag;

      // Transition.  Lane l is in the previous group for u < 2l (window
      // position L_p + u - 2l) and in this one from u = 2l (position u - 2l).
      // Per lane: code base, table row, the first u at which its cells leave
      // the window (vlim) and the u of its band-edge cell (ued), switched at
      // u = 2l; lane 15 stores with the previous group's base before u = 30.
      // New-row values of this lane (row offsets): zero candidate, reset
      // maximum and the F of its first cell.
      const int32_t k0n = ROFF ? k0 + S * L::ONE : L::K0;
      const int32_t hn0n = (ROFF ? k0n + (L::HN0 - L::K0) : L::HN0) + BHO;
      const int32_t ffn = ROFF ? k0n + (F_FIRST - L::K0) : F_FIRST;
      bool fast = false;
      if constexpr (ROFF)
        fast = wave_min((npos_p == L_p && npos_c > XW) ? 1 : 0) != 0;
      // Reduced transition in blocks of FB steps: the addends of a block are
      // read first and the code and table pointers of a lane switch with its
      // row at u = 2l.  BATCHED (groups of at least 2S + FB - 1 steps): the
      // block's row-buffer entries are read first too and its stores issued
      // last -- an entry lane 0 reads at step u was stored by lane S-1 at
      // step u - (L_p - 2S + 1 + pos0_p - pos0_c) or earlier, i.e. in an
      // earlier block; shorter groups read and store them step by step.
      auto xfast = [&](auto batched_c) {
        constexpr bool BATCHED = decltype(batched_c)::value;
        int2* const wp_p = rowbuf + (pos0_p - RAMP + L_p);
        int2* const wp_c = rowbuf + (pos0_c - RAMP);
        const unsigned char* cptr = smem + blo_p + L_p;   // code of step u at cptr[u]
        const unsigned char* tptr = smem + tabo_p;
#pragma unroll
        for (int u0 = 0; u0 < XW; u0 += FB) {
          int32_t tv[FB];
          int2 r2[FB];
#pragma unroll
          for (int v = 0; v < FB; v += 2) {
            const bool sw = lr == ((u0 + v) >> 1);
            cptr = sw ? smem + blo_c : cptr;
            tptr = sw ? smem + tabo_c : tptr;
            tv[v] = *reinterpret_cast<const int32_t*>(tptr + cptr[u0 + v]);
            tv[v + 1] = *reinterpret_cast<const int32_t*>(tptr + cptr[u0 + v + 1]);
          }
          static_assert(FB == 8, "rb_block8");
          if constexpr (BATCHED) rb_block8(rowbuf + rbase + u0, lr, r2);
          int2 wv_[FB];
#pragma unroll
          for (int v = 0; v < FB; ++v) {
            const int u = u0 + v;
            bool sw = false;
            if ((u & 1) == 0) {
              sw = lr == (u >> 1);
              pbest = sw ? st.best : pbest;
              st.best = sw ? hn0n : st.best;
              k0 = sw ? k0n : k0;
            }
            if constexpr (!BATCHED) {
              r2[v] = unspecified_int2();
              if (lr == 0) r2[v] = rowbuf[rbase + u];
            }
            const int32_t hd = st.hu;
            st.hu = __builtin_amdgcn_update_dpp(r2[v].x, st.hout, DPP, 0xF, 0xF, false);
            int32_t eu = __builtin_amdgcn_update_dpp(r2[v].y, st.eout, DPP, 0xF, 0xF, false);
            if (S == 8 && mid) {
              st.hu = r2[v].x;
              eu = r2[v].y;
            }
            const int32_t d = hd + tv[v];
            int32_t e = imax(eu, st.hu) | PREF_E;
            // step 2l - 1: the lane's last cell of the previous group
            if ((u & 1) && lr == ((u + 1) >> 1)) e = L::ENE;
            const int32_t f = sw ? ffn : (imax(st.fl, st.hl) | PREF_F) - L::ONE;
            const int32_t h = imax(imax(d, e), imax(f, k0));
            const int32_t hm = h & ~L::PMASK;
            const int32_t hn = hm - 10 * L::ONE;
            if ((hm >> L::SHIFT) > (st.best >> L::SHIFT)) st.best = hm;
            st.hl = hn;
            st.fl = f;
            st.hout = hn;
            st.eout = e;
            wv_[v] = make_int2(hn, e);
            if constexpr (!BATCHED) {
              if (writer) {
                if (u < RAMP) wp_p[u] = wv_[v];
                else wp_c[u] = wv_[v];
              }
            }
          }
          if constexpr (BATCHED) {
            if (S == 16 && u0 + FB <= RAMP) {
              rb_store8(wp_p + u0, lr, wv_);
            } else if (writer) {
#pragma unroll
              for (int v = 0; v < FB; ++v) {
                if (u0 + v < RAMP) wp_p[u0 + v] = wv_[v];
                else wp_c[u0 + v] = wv_[v];
              }
            }
          }
        }
      };
      if (!SHORTW && fast && L_p >= XW + FB - 1) {
        xfast(std::true_type{});
      } else if (SHORTW && fast) {
        xfast(std::false_type{});
      } else {
        int32_t vlim = npos_p - L_p + 2 * lr;
        int32_t ued = edge_p >= 0 ? edge_p - L_p + 2 * lr : -0x4000;
        const int32_t vlim_c = npos_c + 2 * lr;
        const int32_t ued_c = edge_c >= 0 ? edge_c + 2 * lr : -0x4000;
        const int32_t wb_p = pos0_p - RAMP + L_p, wb_c = pos0_c - RAMP;
        const int32_t cb_p = blo_p + L_p;
        // The addends of XB steps are looked up together (staged codes and
        // table are not written during the transition): at step u a lane
        // reads the new group's buffer and table row once u >= 2l, the
        // previous group's before.  Row-buffer entries are read step by step:
        // with short groups lane 15 writes them only a few steps earlier.
        constexpr int XB = S <= 16 ? 8 : 1;
        int32_t tvb[XB];
#pragma unroll XU
        for (int u = 0; u < XW; ++u) {
          if (u % XB == 0) {
#pragma unroll
            for (int v = 0; v < XB; ++v) {
              const bool in_c = 2 * lr <= u + v;
              tvb[v] = *reinterpret_cast<const int32_t*>(
                  smem + (in_c ? tabo_c : tabo_p) + smem[(in_c ? blo_c : cb_p) + u + v]);
            }
          }
          bool sw = false;
          if ((u & 1) == 0) {
            sw = lr == (u >> 1);
            vlim = sw ? vlim_c : vlim;
            ued = sw ? ued_c : ued;
            pbest = sw ? st.best : pbest;
            st.best = sw ? hn0n : st.best;
            if (ROFF) k0 = sw ? k0n : k0;
          }
          const int32_t tv = tvb[u % XB];
          // Lane 0 is in this group for the whole transition and reads the
          // true row above, also past its window.  There its cells lie right
          // of the matrix (they feed only cells right of it, whose values
          // stay below the in-matrix maximum) or of the band (they feed the
          // band-edge cells through E, which the edge rule replaces); every
          // entry read, up to pos0 + 2S, holds a value of this pair.
          int2 rbv = unspecified_int2();
          if (lr == 0) rbv = rowbuf[rbase + u];
          const int32_t hd = st.hu;
          st.hu = __builtin_amdgcn_update_dpp(rbv.x, st.hout, DPP, 0xF, 0xF, false);
          int32_t eu = __builtin_amdgcn_update_dpp(rbv.y, st.eout, DPP, 0xF, 0xF, false);
          if (S == 8 && mid) {   // first lane of the second segment of a DPP row
            st.hu = rbv.x;
            eu = rbv.y;
          }
          const int32_t d = hd + tv;
          int32_t e = imax(eu, st.hu) | PREF_E;
          if (!ROFF) e -= L::ONE;
          if (u == ued) e = L::ENE;
          // row offsets: step 2l - 1, the lane's last cell of the previous group
          if (ROFF && (u & 1) && lr == ((u + 1) >> 1)) e = L::ENE;
          const int32_t f = sw ? ffn : (imax(st.fl, st.hl) | PREF_F) - L::ONE;
          int32_t h;
          if (WIDE) {
            h = imax(imax(d, e), f);
            if (h < 0) h = L::ZERO_M1;
          } else {
            h = imax(imax(d, e), imax(f, k0));
          }
          const int32_t hm = h & ~L::PMASK;
          const int32_t hn = hm - 10 * L::ONE;
          const int32_t bv = ROFF ? hm : hn;
          const bool valid = u < vlim;
          if (valid && (bv >> L::SHIFT) > (st.best >> L::SHIFT)) st.best = bv;
          st.hl = hn;
          st.fl = f;
          st.hout = hn;
          st.eout = e;
          if (writer && valid) rowbuf[(u < RAMP ? wb_p : wb_c) + u] = make_int2(hn, e);
        }
      }

      // Reduce the previous group's maxima (first maximum: highest score,
      // then lowest row), in row order into the pair's best.
      {
        // row maximum of the previous group without its row offset
        int32_t pb = pbest;
        if constexpr (ROFF) {
          satf |= (live_prev && pbest >= SAT_KEY + BHO) ? 1 : 0;
          pb = pbest - (k0 - L::K0) + S * L::ONE - BHO;
        }
        int32_t r = (pb & ~L::LOW) | (S - 1 - lr);
        if constexpr (S == 16) {   // one DPP row: rotations within it
          r = imax(r, __builtin_amdgcn_update_dpp(r, r, 0x128, 0xF, 0xF, false));   // row_ror:8
          r = imax(r, __builtin_amdgcn_update_dpp(r, r, 0x124, 0xF, 0xF, false));   // row_ror:4
          r = imax(r, __builtin_amdgcn_update_dpp(r, r, 0x122, 0xF, 0xF, false));   // row_ror:2
          r = imax(r, __builtin_amdgcn_update_dpp(r, r, 0x121, 0xF, 0xF, false));   // row_ror:1
        } else {
#pragma unroll
          for (int off = S / 2; off >= 1; off >>= 1) r = imax(r, __shfl_xor(r, off));
        }
        const int32_t wl = (S - 1) - (r & (S - 1));
        const int32_t wbest = __shfl(pb, (lane & ~(S - 1)) | wl);
        if (live_prev && (wbest >> L::SHIFT) > (seg_best >> L::SHIFT)) seg_best = wbest;
      }
      if (gi == smax) break;

      // Stage the next group (or the empty drain group) into the other buffer.
      live_n = window(gi + 1, pos0_n, npos_n, edge_n);
      L_n = (gi + 1 < smax) ? wave_max(imax(XW, npos_n)) : XW;
      if (ROFF && live_n) {
        // entries left of column 1 for the next stripe: H = 0 of the row
        // above it, (gi + 1) * S, with that row's offset (all entries when
        // this stripe's window is empty: it then wrote nothing, lying left
        // of column 1 or right of the matrix, where the next stripe's values
        // do not matter)
        const int2 z = make_int2(L::HN0 + (gi + 1) * S * L::ONE, L::ENN);
        const int32_t hi = npos_c > 0 ? pos0_c : wmax + RAMP + 2;
        for (int32_t x = pos0_n + lr; x < hi; x += S) rowbuf[x] = z;
      }
      {
        const int32_t bo_n = (gi & 1) ? bs0 : bs1;
        if (live_c && live_n && pos0_c == 0 && pos0_n == 0)
          ac_n = stage_shift(gi + 1, bo_n, L_n + STAGE_PAD, (gi & 1) ? bs1 : bs0, L_c + STAGE_PAD);
        else
          ac_n = stage(gi + 1, live_n, pos0_n, bo_n, L_n + STAGE_PAD);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

      // Body of this group: steady while every live window holds, then tail.
      const int32_t wb = pos0_c - RAMP;
      int32_t wsteady = live_c ? (edge_c >= 0 ? npos_c - 1 : npos_c) : 0x7FFFFFFF;
      wsteady = __builtin_amdgcn_readfirstlane(wave_min(wsteady));   // uniform: a scalar bound
      int32_t tau = XW;
      {
        constexpr int U = 8;
        for (; tau + U <= wsteady; tau += U) {
          int32_t tv[U];
          int2 r2[U];
#pragma unroll
          for (int v = 0; v < U; ++v)
            tv[v] = *reinterpret_cast<const int32_t*>(smem + tabo_c + smem[blo_c + tau + v]);
          rb_block8(rowbuf + rbase + tau, lr, r2);
          int2 wv_[U];
          if constexpr (WIDE) {
#pragma unroll
            for (int v = 0; v < U; ++v) {
              dp_step<STEADY_NOWRITE, S, WIDE>(st, tv[v], r2[v], 0, npos_c, edge_c, false,
                                               nullptr, mid);
              wv_[v] = make_int2(st.hout, st.eout);
            }
          } else {
            // First maximum of the block: each step's h (preference cleared)
            // is tagged in bits 13-15 with U - 1 - v by the same and-or that
            // clears them, and HN is that value less tag + 10 << 16 (hidden
            // from the compiler, which would otherwise undo the fusion); one
            // max tree then picks the highest score and, among equal scores,
            // the earliest step, and the running maximum keeps it only on a
            // strictly higher score (the tag bits stay below the score field).
            static_assert(U == 8, "three tag bits");
            int32_t tg[U];
            const int32_t pmask = ~L::PMASK;
#pragma unroll
            for (int v = 0; v < U; ++v) {
              const int32_t tk = ((U - 1 - v) << 13) + 10 * L::ONE;
              const int32_t hd = st.hu;
              st.hu = __builtin_amdgcn_update_dpp(r2[v].x, st.hout, DPP, 0xF, 0xF, false);
              int32_t eu = __builtin_amdgcn_update_dpp(r2[v].y, st.eout, DPP, 0xF, 0xF, false);
              if (S == 8 && mid) {
                st.hu = r2[v].x;
                eu = r2[v].y;
              }
              const int32_t d = hd + tv[v];
              const int32_t e = imax(eu, st.hu) | PREF_E;
              const int32_t f = (imax(st.fl, st.hl) | PREF_F) - L::ONE;
              const int32_t h = imax(imax(d, e), imax(f, k0));
              tg[v] = and_or(h, pmask, (U - 1 - v) << 13);
              const int32_t hn = tg[v] - tk;
              st.hl = hn;
              st.fl = f;
              st.hout = hn;
              st.eout = e;
              wv_[v] = make_int2(hn, e);
            }
            const int32_t m = imax(imax(imax(imax(tg[0], tg[1]), tg[2]), imax(imax(tg[3], tg[4]), tg[5])),
                                   imax(tg[6], tg[7]));
            if ((m >> L::SHIFT) > (st.best >> L::SHIFT)) st.best = m;
          }
          if constexpr (S == 16 && !WIDE) {
            rb_store8(rowbuf + wb + tau, lr, wv_);
          } else if (writer) {
#pragma unroll
            for (int v = 0; v < U; ++v) rowbuf[wb + tau + v] = wv_[v];
          }
        }
      }
#pragma unroll 2
      for (; tau < L_c; ++tau) {
        const int32_t kp = tau - 2 * lr;
        const int32_t tv = *reinterpret_cast<const int32_t*>(smem + tabo_c + smem[blo_c + tau]);
        int2 rbv = unspecified_int2();
        if (lr == 0) rbv = rowbuf[rbase + min(tau, npos_c)];
        const bool wr = writer && (uint32_t)kp < (uint32_t)npos_c;
        dp_step<RAMP_DOWN | (ROFF ? M_BESTH : 0), S, WIDE, ROFF>(
            st, tv, rbv, kp, npos_c, edge_c, wr, rowbuf + (wb + tau), mid, k0);
      }
      live_prev = live_c;
      blo_p = blo_c;
      tabo_p = tabo_c;
      pos0_p = pos0_c;
      npos_p = npos_c;
      edge_p = edge_c;
      L_p = L_c;
    }

    int32_t any_bad = bad | satf;
#pragma unroll
    for (int off = 1; off < S; off <<= 1) any_bad |= __shfl_xor(any_bad, off);
    if (valid && lr == 0) {
      int32_t score = (seg_best >> L::SHIFT) + 11;
      int32_t matches = seg_best & MATCH_MASK;
      if (any_bad || score >= L::SAT) {
        const int32_t slot = atomicAdd(fail_len, 1);
        fail_list[slot] = (int32_t)p;
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

#define CK_PIPE_INST(S_, W_, B_, NT_, SH_)                                                     \
  template __global__ void sw_pipe_kernel<S_, W_, B_, NT_, SH_>(                               \
      const int32_t*, int32_t*, const int32_t*, const int32_t*, const uint8_t*, const int64_t*, \
      const int32_t*, const int32_t*, int32_t, const int8_t*, float, uint32_t, int32_t, int32_t, \
      float*, int32_t*, int32_t*, uint8_t*, int32_t*, int32_t*);
CK_PIPE_INST(16, false, 1, WAVE_BLOCK, false)
CK_PIPE_INST(16, false, 4, WAVE_BLOCK, false)
CK_PIPE_INST(16, false, 1, WAVE_BLOCK, true)
CK_PIPE_INST(16, false, 4, WAVE_BLOCK, true)
CK_PIPE_INST(16, false, 1, 1024, false)
CK_PIPE_INST(8, false, 1, WAVE_BLOCK, false)
CK_PIPE_INST(8, false, 4, WAVE_BLOCK, false)
CK_PIPE_INST(64, false, 1, WAVE_BLOCK, false)
CK_PIPE_INST(64, true, 1, WAVE_BLOCK, false)
#undef CK_PIPE_INST

#define CK_WAVE_INST(S_, W_)                                                                   \
  template __global__ void sw_wave_kernel<S_, W_>(                                             \
      const int32_t*, int32_t*, const int32_t*, const int32_t*, const uint8_t*, const int64_t*, \
      const int32_t*, const int32_t*, int32_t, const int8_t*, float, uint32_t, int32_t, int32_t, \
      float*, int32_t*, int32_t*, uint8_t*, int32_t*, int32_t*);
CK_WAVE_INST(16, false)
CK_WAVE_INST(8, false)
#undef CK_WAVE_INST

}  // namespace ck
