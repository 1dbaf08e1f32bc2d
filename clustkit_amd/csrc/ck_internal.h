// Shared definitions of the MI355X alignment-stage kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ck {

// Scoring constants of the reference (clustkit/csrc/sw_align.c:45-47).
constexpr int32_t GAP_OPEN = -11;
constexpr int32_t GAP_EXTEND = -1;
constexpr int32_t NEG_INF = -1000000;
constexpr int32_t UNKNOWN_SCORE = -4;

// ---------------------------------------------------------------------------
// Packed-key representation used by the batched kernel.
//
// One int32 carries a 16-bit signed score, a 3-bit tie-break preference and
// a 13-bit match count:   key = score << 16 | pref << 13 | matches.
// Ordering keys with a single signed max() orders by score, then by
// preference, then by matches.  Each recurrence gives its candidates the
// preferences that reproduce the reference tie rules:
//   E, F : extension (pref 2 for E, 1 for F) beats opening (pref 0)
//   H    : 0 (pref 4) > diagonal (3) > E (2) > F (1)
// Stored values are pre-shifted so that the additions of the recurrence are
// folded into them:
//   HN = (H - 11) << 16 | m          gap-open candidate, pref 0
//   EN = (E -  1) << 16 | 2<<13 | m  gap-extension candidate of the row below
//   FN = (F -  1) << 16 | 1<<13 | m  gap-extension candidate of the next column
// and the H selection runs in the "minus one" domain, all four candidates
// lowered by 1.  Matches fit 13 bits when the shorter sequence has at most
// 8191 residues; scores fit 16 bits below SAT_LIMIT.  Pairs outside either
// bound are re-run by the int32 kernel.
// ---------------------------------------------------------------------------
constexpr int KEY_SHIFT = 16;
constexpr int32_t KEY_ONE = 1 << KEY_SHIFT;
constexpr int32_t PREF_F = 1 << 13;
constexpr int32_t PREF_E = 2 << 13;
constexpr int32_t PREF_D = 3 << 13;
constexpr int32_t PREF_0 = 4 << 13;
constexpr int32_t PREF_MASK = 7 << 13;
constexpr int32_t MATCH_MASK = (1 << 13) - 1;
constexpr int32_t MAX_FAST_ROWS = 8191;      // shorter sequence bound (13-bit matches)
constexpr int32_t SAT_LIMIT = 32700;         // best score at/above this -> int32 rerun
constexpr int32_t KEY_NEG_SCORE = -30000;    // "minus infinity" score inside keys
constexpr int32_t POISON_SCORE = -16000;     // substitution score of padding residues

constexpr int32_t key_of(int32_t score, int32_t pref, int32_t m) {
  return (int32_t)((uint32_t)score << KEY_SHIFT) + pref + m;
}
constexpr int32_t HN_ZERO = key_of(0 - 11, 0, 0);              // H = 0 outside the band
constexpr int32_t EN_NEG = key_of(KEY_NEG_SCORE, PREF_E, 0);   // E = -inf
constexpr int32_t FN_NEG = key_of(KEY_NEG_SCORE, PREF_F, 0);   // F = -inf
constexpr int32_t EN_EDGE = key_of(-11 - 1, PREF_E, 0);        // E = -11, m = 0
constexpr int32_t K0_SHIFTED = key_of(-1, PREF_0, 0);          // candidate "0"

// Residue codes inside the batched kernel: 0..19 amino acids, 20..30 other
// codes (kept distinct so equal unknown codes still count as matches), 31 =
// padding outside the sequence.  Raw codes > 30 send the pair to the int32
// kernel, which handles every byte value.
constexpr int CODE_POISON = 31;
constexpr int CODE_MAX_FAST = 30;
constexpr int TABLE_DIM = 32;
// Rows of the diagonal-addend table in LDS are TABLE_DIM + 1 words apart, so
// lanes looking up the same column residue for different row residues hit
// different banks (with rows TABLE_DIM apart they share one bank).
constexpr int TABLE_STRIDE = TABLE_DIM + 1;
constexpr int TABLE_BYTES = TABLE_DIM * TABLE_STRIDE * 4;

// Geometry of the batched kernel: a pair takes a segment of 16 lanes (one
// DPP row; four pairs per wavefront) or of 8 lanes (eight pairs per
// wavefront); 256-thread workgroups.
constexpr int WAVE = 64;
constexpr int WAVE_BLOCK = 256;
constexpr int DEFAULT_BIN_SHIFT = 5;
constexpr int MAX_FAST_W = 2 * 255 + 1;      // widest band row held in LDS

// Work-queue bins (ordered by descending estimated cost).
constexpr int NBINS = 4096;

// Device-side counters kept in the workspace.
struct Counters {
  int32_t queue_head;     // next queue slot handed to a wavefront
  int32_t queue_len;      // pairs in the batched queue
  int32_t rerun_len;      // pairs flagged by the 16-bit kernel (wide-score re-run list)
  int32_t skipped;        // pairs removed by the length-ratio test
  int32_t wide_head;      // next re-run slot handed to a wide-score wavefront
  int32_t plain_len;      // pairs for the int32 kernel
  int32_t prep_plain;     // of which routed there directly (rows > 8191, wide band)
  int32_t aux_head;       // next slot of a third queue range (global alignment)
  int32_t err;            // set by a kernel that met an argument it cannot serve
};

// Arguments of the k-mer selection kernels (kmer_score.hip).
// A 20 x 20 substitution table passed by value (kernel argument).
struct KmSub400 {
  int8_t v[400];
};

struct SelectArgs {
  uint32_t* scores;       // [nq][stride]
  int64_t stride;         // per-query stride of scores (entries)
  const int32_t* n_of_q;  // entries per query (device) or nullptr
  int32_t n_fixed;        // entries per query when n_of_q is null
  int32_t min_score;
  int32_t K;
  int32_t* out_ids;       // [nq][K]
  int32_t* out_n;         // [nq]
  int32_t* out_trunc;     // [nq] 1 when more than K passed (or nullptr)
  int32_t mark;           // leave KM_SURV | rank in the selected score words
  // Phase-B histogram offsets (optional): nbins(id) = (len[id] - k + qlen)/w + 1
  int32_t* out_hoff;      // [nq][K+1] or nullptr
  const int32_t* id_len;  // target lengths
  const int32_t* qlen;
  int32_t k, w;
  int32_t* ghist;         // [nq][KM_SAT+1], zero on entry, left zero
  int4* plan;             // [nq] {n, cut, take, total}
  int32_t* bcnt;          // [nq][B][2] above / ties per slice
};

// Per-pair orientation and geometry shared by both kernels.
struct PairGeom {
  int64_t a_off, b_off;   // flat offsets: a = rows (shorter), b = columns
  int32_t la, lb;         // la <= lb
  int32_t bw;             // effective half band
  int32_t hint;           // diagonal offset, oriented
};

__device__ __forceinline__ bool length_ratio_fails(int32_t li, int32_t lj, float t) {
  int32_t s = li < lj ? li : lj;
  int32_t l = li > lj ? li : lj;
  return l > 0 && ((float)s / (float)l) < t;
}

// Orientation rules of batch_sw_align_c (sw_align.c:237-256): the shorter
// sequence indexes rows; hints beyond +-shorter are dropped; a swap negates
// the hint.  Band rule of sw_align.c:55-58.
__device__ __forceinline__ PairGeom pair_geom(const int32_t* pairs, int64_t p,
                                              const int64_t* offsets, const int32_t* lengths,
                                              const int32_t* hints, int32_t band) {
  int32_t x = pairs[2 * p], y = pairs[2 * p + 1];
  int32_t lx = lengths[x], ly = lengths[y];
  int32_t h = hints ? hints[p] : 0;
  int32_t sh = lx < ly ? lx : ly;
  if (h > sh || h < -sh) h = 0;
  PairGeom g;
  if (lx <= ly) {
    g.a_off = offsets[x]; g.b_off = offsets[y]; g.la = lx; g.lb = ly; g.hint = h;
  } else {
    g.a_off = offsets[y]; g.b_off = offsets[x]; g.la = ly; g.lb = lx; g.hint = -h;
  }
  int32_t longest = g.lb;
  g.bw = (band <= 0 || longest <= 50) ? longest : band;
  return g;
}

// Wave-uniform work fetch of a persistent loop: lane 0 adds n to *head and
// the wavefront reads lane 0's old value as a scalar.  Every lane executes
// the atomic (the others add 0), so no lane-dependent branch surrounds it:
// with a lane-0-only atomic under a branch and a per-lane constant for the
// other lanes, the compiler may turn the persistent loop into two nested
// loops split by lane, in which lanes 1..63 never fetch again.
__device__ __forceinline__ int32_t wave_fetch(int32_t* head, int32_t n) {
  const int lane = threadIdx.x & (WAVE - 1);
  return __builtin_amdgcn_readfirstlane(atomicAdd(head, lane == 0 ? n : 0));
}

}  // namespace ck
