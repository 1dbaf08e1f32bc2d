// k-mer inverted-index scoring of query batches (device side).
//
// Semantics: clustkit/csrc/kmer_score.c:198-472 (score_query_full,
// batch_score_queries_c):
//   Phase A   count[q][t] = hits of the query's k-mers on target t, a query
//             k-mer counting only when all its residues are < alpha and its
//             index frequency is <= freq_thresh; saturating at 32767.
//   survivors targets with count >= min_total_hits; when more than
//             phase_a_topk pass, the top phase_a_topk by (count desc, id asc).
//   Phase B   per survivor, histogram of diagonal bins (tpos - qpos + qlen)/w
//             over its hits; best bin = highest count, lowest bin on ties;
//             kept when >= min_diag_hits; hint = bin*w - qlen + w/2.
//   output    top max_cands of the kept targets by (score desc, id asc);
//             ascending id when no more than max_cands are kept.  With
//             min_diag_hits <= 1 the Phase-A counts are ranked directly.
//
// Device layout.  Postings are split into int32 target ids and int32
// positions (the two halves of the reference's packed int64 entries).  For
// dense indexes (k = 3) the targets are cut into tiles of TILE ids and a
// per-(k-mer, tile) offset table lets one workgroup walk only the part of
// each posting list that falls in its tile, counting hits with LDS atomics;
// sparse indexes (k = 5) count with global atomics.  Selections are exact:
// a 32768-bin histogram of the (saturated) scores gives the cut value, and
// an ordered ballot scan over ids resolves ties at the cut by ascending id.
#include <type_traits>

#include "ck_internal.h"

namespace ck {

constexpr int KM_TILE = 8192;           // targets per Phase-A/B tile (32 KB of LDS counters)
constexpr int KM_SAT = 32767;
constexpr uint32_t KM_SURV = 0x80000000u;

// Query k-mer codes: qcode[qoff + p] = code, or -1 when the window holds a
// residue >= alpha or the code's index frequency exceeds the cap.
__global__ __launch_bounds__(256) void km_qcodes_kernel(
    const uint8_t* __restrict__ qflat, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, int32_t nq, int32_t k, int32_t alpha,
    const int32_t* __restrict__ freqs, int32_t freq_thresh, int32_t* __restrict__ qcode) {
  for (int q = blockIdx.y; q < nq; q += gridDim.y) {
    const int32_t nk = qlen[q] - k + 1;
    const uint8_t* s = qflat + qoff[q];
    for (int32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < nk;
         p += gridDim.x * blockDim.x) {
      int64_t c = 0;
      bool ok = true;
      for (int j = 0; j < k; ++j) {
        const int32_t r = s[p + j];
        ok &= r < alpha;
        c = c * alpha + (r < alpha ? r : 0);
      }
      if (ok && freqs[c] > freq_thresh) ok = false;
      qcode[qoff[q] + p] = ok ? (int32_t)c : -1;
    }
  }
}

__global__ __launch_bounds__(256) void km_fill_kernel(uint32_t* __restrict__ p, int64_t n,
                                                      uint32_t v) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n;
       x += (int64_t)gridDim.x * blockDim.x)
    p[x] = v;
}

// Phase A, tiled: grid (tiles, queries, splits).  counts[q*nd + t] = raw hit
// count.  With one split the workgroup stores its tile's counts; with more
// (small batches: more workgroups per query) each split takes a slice of the
// query's k-mers and adds its counts into counts, which must start at zero.
__global__ __launch_bounds__(256) void km_phase_a_tiled_kernel(
    const int32_t* __restrict__ qcode, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, int32_t k, const int64_t* __restrict__ tile_off,
    int32_t ntiles, const int32_t* __restrict__ ptid, int32_t nd,
    uint32_t* __restrict__ counts) {
  __shared__ uint32_t cnt[KM_TILE];
  const int j = blockIdx.x, q = blockIdx.y;
  const int32_t t0 = j * KM_TILE;
  for (int x = threadIdx.x; x < KM_TILE; x += blockDim.x) cnt[x] = 0;
  __syncthreads();
  const int32_t nk = qlen[q] - k + 1;
  const int32_t* qc = qcode + qoff[q];
  const int32_t span = (max(nk, 0) + (int32_t)gridDim.z - 1) / (int32_t)gridDim.z;
  const int32_t p_lo = (int32_t)blockIdx.z * span, p_hi = min(nk, p_lo + span);
  // Each wavefront takes 64 k-mers at a time: every lane fetches one
  // k-mer's tile slice, then the wavefront walks the slices one after
  // another with the bounds broadcast from their lane (readlane), so the
  // slice lookups of 64 k-mers are in flight together.
  // The workgroup's k-mers are split evenly over its wavefronts (a small
  // batch gives a workgroup few k-mers), 64 at a time within a wavefront.
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, nw = blockDim.x / WAVE;
  const int32_t cw = (max(p_hi - p_lo, 0) + nw - 1) / nw;
  const int32_t w_lo = p_lo + wv * cw, w_hi = min(p_hi, w_lo + cw);
  for (int32_t p0 = w_lo; p0 < w_hi; p0 += WAVE) {
    const int32_t nc = min(WAVE, w_hi - p0);
    int64_t b = 0, e = 0;
    if (lane < nc) {
      const int32_t c = qc[p0 + lane];
      if (c >= 0) {
        const int64_t* to = tile_off + (int64_t)c * (ntiles + 1) + j;
        b = to[0];
        e = to[1];
      }
    }
    const int32_t len = (int32_t)(e - b);
    for (int32_t i = 0; i < nc; ++i) {
      const int64_t bi = ((int64_t)__builtin_amdgcn_readlane((int32_t)(b >> 32), i) << 32) |
                         (uint32_t)__builtin_amdgcn_readlane((int32_t)b, i);
      const int32_t li = __builtin_amdgcn_readlane(len, i);
      const int32_t* pt = ptid + bi;
      // Four posting loads in flight per lane before the counter updates.
      for (int32_t u = lane; u < li; u += 4 * WAVE) {
        int32_t x[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) x[r] = u + r * WAVE < li ? pt[u + r * WAVE] : -1;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (x[r] >= 0) atomicAdd(&cnt[x[r] - t0], 1u);
      }
    }
  }
  __syncthreads();
  const int32_t n = min(KM_TILE, nd - t0);
  uint32_t* out = counts + (int64_t)q * nd + t0;
  if (gridDim.z == 1) {
    for (int x = threadIdx.x; x < n; x += blockDim.x) out[x] = cnt[x];
  } else {
    for (int x = threadIdx.x; x < n; x += blockDim.x)
      if (cnt[x]) atomicAdd(&out[x], cnt[x]);
  }
}

// Phase A, global atomics: grid (chunks, queries) over the query's k-mers.
// BLOSUM-similar Phase A (kmer_index.py:1186-1297, k = 5, alphabet 20): for
// every query 5-mer q with all residues < 20 and every 5-mer a with column
// score S = sum_j sub[q_j][a_j] >= T and 0 < freq(a) <= ft, each posting of
// a adds S to its target's count.  A workgroup takes pos_per_block query
// positions; its threads split the 8000 three-residue prefixes (a0, a1, a2)
// of each position, drop those whose best completion (the row maxima of
// sub for q3 and q4) stays below T, and walk a3, a4 below the rest (the same
// bound at a3).  The bounds only skip codes that cannot reach T, so the
// codes counted are exactly those with S >= T.  T >= 1: counts only grow.
__global__ __launch_bounds__(256) void km_similar_kernel(
    const uint8_t* __restrict__ qflat, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, const int64_t* __restrict__ kmer_offsets,
    const int32_t* __restrict__ post_tid, const int32_t* __restrict__ kmer_freqs,
    int32_t num_db, KmSub400 sub, int32_t T, int32_t ft, int32_t pos_per_block,
    uint32_t* __restrict__ counts) {
  __shared__ int32_t S[400];
  __shared__ int32_t mx[20];
  for (int x = threadIdx.x; x < 400; x += blockDim.x) S[x] = sub.v[x];
  __syncthreads();
  if (threadIdx.x < 20) {
    int32_t m = S[threadIdx.x * 20];
    for (int j = 1; j < 20; ++j) m = max(m, S[threadIdx.x * 20 + j]);
    mx[threadIdx.x] = m;
  }
  __syncthreads();
  const int q = blockIdx.y;
  const int32_t nk = qlen[q] - 4;
  const int32_t p0 = (int32_t)blockIdx.x * pos_per_block;
  const int32_t p1 = min(nk, p0 + pos_per_block);
  const uint8_t* qs = qflat + qoff[q];
  uint32_t* cnt = counts + (int64_t)q * num_db;
  for (int32_t p = p0; p < p1; ++p) {
    const int32_t r0 = qs[p], r1 = qs[p + 1], r2 = qs[p + 2], r3 = qs[p + 3], r4 = qs[p + 4];
    if (r0 >= 20 || r1 >= 20 || r2 >= 20 || r3 >= 20 || r4 >= 20) continue;
    const int32_t m4 = mx[r4], m34 = mx[r3] + m4;
    for (int32_t x = threadIdx.x; x < 8000; x += blockDim.x) {
      const int32_t a0 = x / 400, a1 = (x / 20) % 20, a2 = x % 20;
      const int32_t s2 = S[r0 * 20 + a0] + S[r1 * 20 + a1] + S[r2 * 20 + a2];
      if (s2 + m34 < T) continue;
      for (int32_t a3 = 0; a3 < 20; ++a3) {
        const int32_t s3 = s2 + S[r3 * 20 + a3];
        if (s3 + m4 < T) continue;
        for (int32_t a4 = 0; a4 < 20; ++a4) {
          const int32_t s4 = s3 + S[r4 * 20 + a4];
          if (s4 < T) continue;
          const int32_t code = x * 400 + a3 * 20 + a4;
          const int32_t f = kmer_freqs[code];
          if (f == 0 || f > ft) continue;
          const int64_t e = kmer_offsets[code + 1];
          for (int64_t h = kmer_offsets[code]; h < e; ++h)
            atomicAdd(&cnt[post_tid[h]], (uint32_t)s4);
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void km_phase_a_global_kernel(
    const int32_t* __restrict__ qcode, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, int32_t k, const int64_t* __restrict__ koff,
    const int32_t* __restrict__ ptid, int32_t nd, uint32_t* __restrict__ counts) {
  const int q = blockIdx.y;
  const int32_t nk = qlen[q] - k + 1;
  const int32_t* qc = qcode + qoff[q];
  uint32_t* cq = counts + (int64_t)q * nd;
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const int nw = blockDim.x / WAVE;
  for (int32_t p = blockIdx.x * nw + wv; p < nk; p += gridDim.x * nw) {
    const int32_t c = qc[p];
    if (c < 0) continue;
    for (int64_t h = koff[c] + lane; h < koff[c + 1]; h += WAVE) atomicAdd(&cq[ptid[h]], 1u);
  }
}

// Exact top-K of a dense score array by (score desc, id asc), emitted in
// ascending id order, spread over several workgroups per query:
//   km_sel_hist_kernel   grid (B, queries): histogram of saturated scores of
//                        each workgroup's id slice -> per-query histogram
//   km_sel_cut_kernel    grid (queries): cut value and number of ties taken
//                        (resets the per-query histogram for the next use)
//   km_sel_count_kernel  grid (B, queries): entries above the cut and ties at
//                        the cut per slice
//   km_sel_emit_kernel   grid (B, queries): ranks in ascending id (slice
//                        offsets from the counts of the slices before it)
//   km_sel_hoff_kernel   grid (queries): Phase-B histogram offsets of the
//                        selected ids (optional)
// Scores are read as min(score & 0x7FFFFFFF, 32767); only ids with
// score >= min_score take part.  Optionally the selection leaves each
// selected id's rank in its score word (KM_SURV | rank) for Phase B.

constexpr int SEL_LDS_BINS = 2048;              // scores below this count in LDS

__device__ __forceinline__ int32_t sel_n(const SelectArgs& a, int q) {
  int32_t n = a.n_of_q ? a.n_of_q[q] : a.n_fixed;
  if (a.qlen && a.qlen[q] < a.k) n = 0;   // no k-mer in the query: no candidate
  return n;
}

__device__ __forceinline__ void sel_slice(int32_t n, int32_t* lo, int32_t* hi) {
  const int32_t span = (n + (int32_t)gridDim.x - 1) / (int32_t)gridDim.x;
  *lo = min(n, (int32_t)blockIdx.x * span);
  *hi = min(n, *lo + span);
}

__device__ __forceinline__ int32_t sel_score(const uint32_t* sc, int32_t t) {
  return (int32_t)min(sc[t] & 0x7FFFFFFFu, (uint32_t)KM_SAT);
}

// Exclusive prefix over the block's threads (thread order) and the block
// total; nw = wavefronts in the block (<= 16).  All threads must call it.
__device__ __forceinline__ int32_t block_excl_scan(int32_t x, int32_t* ws, int32_t* total) {
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const int nw = blockDim.x / WAVE;
  int32_t incl = x;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const int32_t y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  if (lane == WAVE - 1) ws[wv] = incl;
  __syncthreads();
  if (wv == 0) {
    int32_t v = lane < nw ? ws[lane] : 0;
    int32_t inc = v;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const int32_t y = __shfl_up(inc, off);
      if (lane >= off) inc += y;
    }
    if (lane < nw) ws[16 + lane] = inc - v;
    if (lane == nw - 1) ws[15] = inc;
  }
  __syncthreads();
  const int32_t res = ws[16 + wv] + incl - x;
  *total = ws[15];
  __syncthreads();
  return res;
}

__global__ __launch_bounds__(256) void km_sel_hist_kernel(SelectArgs a) {
  __shared__ int32_t lh[SEL_LDS_BINS];
  const int q = blockIdx.y;
  const uint32_t* sc = a.scores + (int64_t)q * a.stride;
  int32_t* gh = a.ghist + (int64_t)q * (KM_SAT + 1);
  int32_t lo, hi;
  sel_slice(sel_n(a, q), &lo, &hi);
  for (int x = threadIdx.x; x < SEL_LDS_BINS; x += blockDim.x) lh[x] = 0;
  __syncthreads();
  for (int32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
    const int32_t v = sel_score(sc, t);
    if (v < a.min_score) continue;
    if (v < SEL_LDS_BINS) atomicAdd(&lh[v], 1);
    else atomicAdd(&gh[v], 1);
  }
  __syncthreads();
  for (int x = threadIdx.x; x < SEL_LDS_BINS; x += blockDim.x)
    if (lh[x]) atomicAdd(&gh[x], lh[x]);
}

__global__ __launch_bounds__(1024) void km_sel_cut_kernel(SelectArgs a) {
  __shared__ int32_t part[1024];
  __shared__ int32_t s_cut, s_take;
  const int q = blockIdx.x, tid = threadIdx.x;
  int32_t* gh = a.ghist + (int64_t)q * (KM_SAT + 1);
  // each thread owns 32 consecutive bins (KM_SAT + 1 = 32768 = 1024 x 32)
  int32_t h[32];
  int32_t local = 0;
#pragma unroll
  for (int b = 0; b < 32; ++b) {
    h[b] = gh[tid * 32 + b];
    local += h[b];
  }
#pragma unroll
  for (int b = 0; b < 32; ++b)
    if (h[b]) gh[tid * 32 + b] = 0;              // leave the histogram zero
  part[tid] = local;
  if (tid == 0) { s_cut = -1; s_take = 0; }
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {      // inclusive suffix sums
    const int32_t add = tid + off < 1024 ? part[tid + off] : 0;
    __syncthreads();
    part[tid] += add;
    __syncthreads();
  }
  const int32_t total = part[0];
  if (total > a.K) {
    int32_t above = part[tid] - local;
    if (above < a.K && part[tid] >= a.K) {
      for (int b = 31; b >= 0; --b) {
        if (above + h[b] >= a.K) { s_cut = tid * 32 + b; s_take = a.K - above; break; }
        above += h[b];
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    const bool trunc = total > a.K;
    a.plan[q] = make_int4(sel_n(a, q), trunc ? s_cut : a.min_score, trunc ? s_take : 0, total);
    if (a.out_trunc) a.out_trunc[q] = trunc ? 1 : 0;
    a.out_n[q] = trunc ? a.K : total;
  }
}

// above: score > cut (all passing scores when not truncated); tie: == cut
__device__ __forceinline__ void sel_class(int32_t v, const int4& pl, int32_t K, int32_t min_score,
                                          bool* above, bool* tie) {
  const bool trunc = pl.w > K;
  const bool in = v >= min_score;
  *above = trunc ? (in && v > pl.y) : in;
  *tie = trunc && in && v == pl.y;
}

__global__ __launch_bounds__(256) void km_sel_count_kernel(SelectArgs a) {
  __shared__ int32_t ws[32];
  const int q = blockIdx.y;
  const int4 pl = a.plan[q];
  const uint32_t* sc = a.scores + (int64_t)q * a.stride;
  int32_t lo, hi;
  sel_slice(pl.x, &lo, &hi);
  int32_t na = 0, nt = 0;
  for (int32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
    bool ab, ti;
    sel_class(sel_score(sc, t), pl, a.K, a.min_score, &ab, &ti);
    na += ab;
    nt += ti;
  }
  int32_t ta, tt;
  block_excl_scan(na, ws, &ta);
  block_excl_scan(nt, ws, &tt);
  if (threadIdx.x == 0) {
    int32_t* bc = a.bcnt + ((int64_t)q * gridDim.x + blockIdx.x) * 2;
    bc[0] = ta;
    bc[1] = tt;
  }
}

constexpr int SEL_PER = 16;                     // consecutive ids per thread per chunk

__global__ __launch_bounds__(256) void km_sel_emit_kernel(SelectArgs a) {
  __shared__ int32_t ws[32];
  __shared__ int32_t s_ab, s_tb;
  // one chunk of scores, read coalesced and handed to the thread that owns
  // 16 consecutive ids (row of 17 words: conflict-free reads)
  __shared__ uint32_t stg[256 * (SEL_PER + 1)];
  const int q = blockIdx.y, tid = threadIdx.x;
  const int4 pl = a.plan[q];
  const int32_t take = pl.w > a.K ? pl.z : 0;
  uint32_t* sc = a.scores + (int64_t)q * a.stride;
  int32_t lo, hi;
  sel_slice(pl.x, &lo, &hi);
  if (tid == 0) {
    const int32_t* bc = a.bcnt + (int64_t)q * gridDim.x * 2;
    int32_t ab = 0, tb = 0;
    for (int b = 0; b < (int)blockIdx.x; ++b) { ab += bc[2 * b]; tb += bc[2 * b + 1]; }
    s_ab = ab;
    s_tb = tb;
  }
  __syncthreads();
  int32_t above_before = s_ab, ties_before = s_tb;
  int32_t* oid = a.out_ids + (int64_t)q * a.K;
  const int32_t CH = blockDim.x * SEL_PER;
  for (int32_t c0 = lo; c0 < hi; c0 += CH) {
    const int32_t b = c0 + tid * SEL_PER;
#pragma unroll
    for (int u = 0; u < SEL_PER; ++u) {
      const int32_t r = u * (int32_t)blockDim.x + tid;
      stg[(r / SEL_PER) * (SEL_PER + 1) + r % SEL_PER] = c0 + r < hi ? sc[c0 + r] : 0u;
    }
    __syncthreads();
    uint32_t abits = 0, tbits = 0;
#pragma unroll
    for (int u = 0; u < SEL_PER; ++u) {
      if (b + u < hi) {
        bool ab, ti;
        const int32_t v = (int32_t)min(stg[tid * (SEL_PER + 1) + u] & 0x7FFFFFFFu, (uint32_t)KM_SAT);
        sel_class(v, pl, a.K, a.min_score, &ab, &ti);
        if (ab) abits |= 1u << u;
        if (ti) tbits |= 1u << u;
      }
    }
    int32_t tot;
    const int32_t pre = block_excl_scan(__popc(abits) | (__popc(tbits) << 16), ws, &tot);
    int32_t ab = above_before + (pre & 0xFFFF);
    int32_t tb = ties_before + (pre >> 16);
#pragma unroll
    for (int u = 0; u < SEL_PER; ++u) {
      const bool isa = (abits >> u) & 1u, ist = (tbits >> u) & 1u;
      if (isa || (ist && tb < take)) {
        const int32_t rank = ab + min(take, tb);
        oid[rank] = b + u;
        if (a.mark) sc[b + u] = KM_SURV | (uint32_t)rank;   // survivor slot for Phase B
      }
      ab += isa ? 1 : 0;
      tb += ist ? 1 : 0;
    }
    above_before += tot & 0xFFFF;
    ties_before += tot >> 16;
    __syncthreads();   // every thread has read its scores before the next chunk's
  }
}

__global__ __launch_bounds__(1024) void km_sel_hoff_kernel(SelectArgs a) {
  __shared__ int32_t ws[32];
  const int q = blockIdx.x, tid = threadIdx.x;
  const int32_t cnt = a.out_n[q];
  const int32_t* oid = a.out_ids + (int64_t)q * a.K;
  int32_t* ohoff = a.out_hoff + (int64_t)q * (a.K + 1);
  const int32_t ql = a.qlen ? a.qlen[q] : 0;
  int32_t base = 0;
  for (int32_t r0 = 0; r0 < cnt; r0 += blockDim.x) {
    const int32_t r = r0 + tid;
    const int32_t nb = r < cnt ? (a.id_len[oid[r]] - a.k + ql) / a.w + 1 : 0;
    int32_t tot;
    const int32_t pre = block_excl_scan(nb, ws, &tot);
    if (r < cnt) ohoff[r] = base + pre;
    base += tot;
  }
  if (tid == 0) ohoff[cnt] = base;
}

// Phase B, tiled: hits on survivor targets of the tile -> diagonal-bin
// histograms.  grid (tiles, queries).  The selection left each survivor's
// slot (its index in the ascending survivor list) in counts as
// KM_SURV | slot; SLOT16 stages the tile's slots in LDS as uint16 (survivor
// lists below 65535), otherwise they are read from counts.
template <bool SLOT16>
__global__ __launch_bounds__(256) void km_phase_b_tiled_kernel(
    const int32_t* __restrict__ qcode, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, int32_t k, const int64_t* __restrict__ tile_off,
    int32_t ntiles, const int32_t* __restrict__ ptid, const int32_t* __restrict__ ppos,
    int32_t nd, const uint32_t* __restrict__ counts, const int32_t* __restrict__ hoff,
    int32_t K, int32_t w, int32_t* __restrict__ hist, int64_t hist_stride) {
  __shared__ uint16_t slot[SLOT16 ? KM_TILE : 1];
  const int j = blockIdx.x, q = blockIdx.y;
  const int32_t t0 = j * KM_TILE;
  const int32_t n = min(KM_TILE, nd - t0);
  const uint32_t* cq = counts + (int64_t)q * nd + t0;
  bool any = false;
  for (int x = threadIdx.x; x < KM_TILE; x += blockDim.x) {
    const uint32_t v = x < n ? cq[x] : 0u;
    const bool s = (v & KM_SURV) != 0;
    any |= s;
    if (SLOT16) slot[x] = s ? (uint16_t)(v & 0xFFFFu) : (uint16_t)0xFFFFu;
  }
  if (!__syncthreads_or(any)) return;
  const int32_t ql = qlen[q];
  const int32_t nk = ql - k + 1;
  const int32_t* qc = qcode + qoff[q];
  const int32_t* ho = hoff + (int64_t)q * (K + 1);
  int32_t* hq = hist + (int64_t)q * hist_stride;
  const int32_t span = (max(nk, 0) + (int32_t)gridDim.z - 1) / (int32_t)gridDim.z;
  const int32_t p_lo = (int32_t)blockIdx.z * span, p_hi = min(nk, p_lo + span);
  // 64 k-mers per wavefront at a time, slice bounds broadcast by readlane
  // (as in km_phase_a_tiled_kernel).
  // The workgroup's k-mers are split evenly over its wavefronts (a small
  // batch gives a workgroup few k-mers), 64 at a time within a wavefront.
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE, nw = blockDim.x / WAVE;
  const int32_t cw = (max(p_hi - p_lo, 0) + nw - 1) / nw;
  const int32_t w_lo = p_lo + wv * cw, w_hi = min(p_hi, w_lo + cw);
  for (int32_t p0 = w_lo; p0 < w_hi; p0 += WAVE) {
    const int32_t nc = min(WAVE, w_hi - p0);
    int64_t b = 0, e = 0;
    if (lane < nc) {
      const int32_t c = qc[p0 + lane];
      if (c >= 0) {
        const int64_t* to = tile_off + (int64_t)c * (ntiles + 1) + j;
        b = to[0];
        e = to[1];
      }
    }
    const int32_t len = (int32_t)(e - b);
    for (int32_t i = 0; i < nc; ++i) {
      const int64_t bi = ((int64_t)__builtin_amdgcn_readlane((int32_t)(b >> 32), i) << 32) |
                         (uint32_t)__builtin_amdgcn_readlane((int32_t)b, i);
      const int32_t li = __builtin_amdgcn_readlane(len, i);
      const int32_t dp = ql - (p0 + i);
      const int32_t* pt = ptid + bi;
      const int32_t* pp = ppos + bi;
      // Four postings per lane in flight: target ids, then the positions
      // of those that hit a survivor, then the histogram updates.
      for (int32_t u = lane; u < li; u += 4 * WAVE) {
        int32_t sl[4], ps[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) sl[r] = u + r * WAVE < li ? pt[u + r * WAVE] - t0 : -1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (sl[r] < 0) continue;
          if (SLOT16) {
            const int32_t v = slot[sl[r]];
            sl[r] = v == 0xFFFF ? -1 : v;
          } else {
            const uint32_t v = cq[sl[r]];
            sl[r] = (v & KM_SURV) ? (int32_t)(v & 0x7FFFFFFFu) : -1;
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) ps[r] = sl[r] >= 0 ? pp[u + r * WAVE] : 0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (sl[r] >= 0) atomicAdd(&hq[ho[sl[r]] + (ps[r] + dp) / w], 1);
      }
    }
  }
}
template __global__ void km_phase_b_tiled_kernel<true>(
    const int32_t*, const int64_t*, const int32_t*, int32_t, const int64_t*, int32_t,
    const int32_t*, const int32_t*, int32_t, const uint32_t*, const int32_t*, int32_t, int32_t,
    int32_t*, int64_t);
template __global__ void km_phase_b_tiled_kernel<false>(
    const int32_t*, const int64_t*, const int32_t*, int32_t, const int64_t*, int32_t,
    const int32_t*, const int32_t*, int32_t, const uint32_t*, const int32_t*, int32_t, int32_t,
    int32_t*, int64_t);

// Phase B with the slots read from counts (sparse indexes).
__global__ __launch_bounds__(256) void km_phase_b_global_kernel(
    const int32_t* __restrict__ qcode, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, int32_t k, const int64_t* __restrict__ koff,
    const int32_t* __restrict__ ptid, const int32_t* __restrict__ ppos, int32_t nd,
    const uint32_t* __restrict__ counts, const int32_t* __restrict__ hoff, int32_t K,
    int32_t w, int32_t* __restrict__ hist, int64_t hist_stride) {
  const int q = blockIdx.y;
  const int32_t ql = qlen[q];
  const int32_t nk = ql - k + 1;
  const int32_t* qc = qcode + qoff[q];
  const uint32_t* cq = counts + (int64_t)q * nd;
  const int32_t* ho = hoff + (int64_t)q * (K + 1);
  int32_t* hq = hist + (int64_t)q * hist_stride;
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const int nw = blockDim.x / WAVE;
  for (int32_t p = blockIdx.x * nw + wv; p < nk; p += gridDim.x * nw) {
    const int32_t c = qc[p];
    if (c < 0) continue;
    for (int64_t h = koff[c] + lane; h < koff[c + 1]; h += WAVE) {
      const uint32_t v = cq[ptid[h]];
      if (!(v & KM_SURV)) continue;
      atomicAdd(&hq[ho[v & 0x7FFFFFFFu] + (ppos[h] - p + ql) / w], 1);
    }
  }
}

// Best diagonal bin per survivor (first maximum); resets the bins it reads
// so the histogram pool is zero for the next batch.  16 lanes per survivor
// read its bins side by side and reduce (count, first bin); grid (chunks,
// queries), 16 survivors per 256-thread workgroup.
__global__ __launch_bounds__(256) void km_phase_b_best_kernel(
    const int32_t* __restrict__ surv_n, const int32_t* __restrict__ hoff, int32_t K,
    int32_t* __restrict__ hist, int64_t hist_stride, uint32_t* __restrict__ best,
    int32_t* __restrict__ best_bin) {
  constexpr int G = 16;
  const int q = blockIdx.y;
  const int32_t ns = surv_n[q];
  const int32_t* ho = hoff + (int64_t)q * (K + 1);
  int32_t* hq = hist + (int64_t)q * hist_stride;
  const int sub = threadIdx.x & (G - 1);
  const int32_t per_block = blockDim.x / G;
  for (int32_t s0 = blockIdx.x * per_block; s0 < ns; s0 += gridDim.x * per_block) {
    const int32_t s = s0 + (int32_t)(threadIdx.x / G);
    // key = count << 32 | (2^32 - 1 - bin): the largest key is the highest
    // count at its first bin.
    uint64_t key = 0;
    if (s < ns) {
      const int32_t b = ho[s], e = ho[s + 1];
      for (int32_t x = b + sub; x < e; x += G) {
        const int32_t v = hq[x];
        if (v) hq[x] = 0;
        const uint64_t c = ((uint64_t)(uint32_t)v << 32) | (uint32_t)(0xFFFFFFFFu - (uint32_t)(x - b));
        key = c > key ? c : key;
      }
    }
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) {
      const uint32_t lo = __shfl_xor((uint32_t)key, off), hi = __shfl_xor((uint32_t)(key >> 32), off);
      const uint64_t o = ((uint64_t)hi << 32) | lo;
      key = o > key ? o : key;
    }
    if (s < ns && sub == 0) {
      const uint32_t bc = (uint32_t)(key >> 32);
      const int32_t bb = key == 0 ? 0 : (int32_t)(0xFFFFFFFFu - (uint32_t)key);
      best[(int64_t)q * K + s] = bc;
      best_bin[(int64_t)q * K + s] = bb;
    }
  }
}

// Phase B, target-major (dense code spaces with the database sequences in
// HBM).  A posting of code c at (t, tpos) is exactly a window of target t at
// tpos whose k residues are all below alpha and whose code is c, so walking
// a survivor's windows and looking each code up among the query's usable
// k-mer positions yields the same (qpos, tpos) hits as walking the query's
// posting lists, while reading only the survivors' residues.
// grid (chunks, queries), 1024 threads.  The workgroup sorts its query's
// usable k-mer positions by code into LDS (counting sort; tab[c] = start |
// end << 16 of code c's run in pos) and stages its survivors' lengths and
// offsets; each wavefront takes a contiguous range of equal total cost
// (length + 64 per survivor), four survivors at a time with their first
// residue chunks loaded together: lane l holds
// bytes [8 l, 8 l + 12) of a 512-byte chunk (one 8-byte and one 4-byte load
// from an 8-aligned base) and scores the 8 windows starting in its first 8
// bytes (rolling codes, a SWAR mask of residues >= alpha).  Hits go to an LDS
// histogram of the diagonal bins (tpos - qpos + qlen)/w in passes of
// KM_TGT_BINS bins; best = highest count at its first bin.  Writes best and
// best_bin as km_phase_b_best_kernel does; the global histogram pool is not
// used.
constexpr int KM_TGT_BINS = 512;
constexpr int KM_TGT_PER_WG = 256;        // survivors per workgroup at most
constexpr int KM_TGT_CHUNK = 8 * WAVE;   // residue bytes per wavefront load

// Bytes [a, a + 12) of the database (flat + a 8-byte aligned) as three
// dwords; bytes outside the buffer read 0 (they only fill windows outside a
// sequence, which are not scored).
__device__ __forceinline__ void km_load12(const uint8_t* __restrict__ flat, int64_t flat_len,
                                          int64_t a, uint32_t (&d)[3]) {
  if (a >= 0 && a + 12 <= flat_len) {
    const uint2 v = *reinterpret_cast<const uint2*>(flat + a);
    d[0] = v.x;
    d[1] = v.y;
    d[2] = *reinterpret_cast<const uint32_t*>(flat + a + 8);
    return;
  }
#pragma unroll
  for (int x = 0; x < 3; ++x) d[x] = 0;
  for (int b = 0; b < 12; ++b)
    if (a + b >= 0 && a + b < flat_len) d[b >> 2] |= (uint32_t)flat[a + b] << (8 * (b & 3));
}

// Scale of diagonal numbers in the hit path: 128 in the FDIV kernels.
template <bool FDIV>
constexpr int32_t KM_NS = FDIV ? 128 : 1;

struct KmTgtQuery {
  const uint32_t* tab;   // [ncodes] 0 (absent), qlen - qpos (one position; times
                         // 128 in the FDIV kernels), bit 31 | a | b << 16 (two
                         // positions) or bit 31 | bit 15 | start | end << 16
                         // (a run in pos); tab[-1] is empty (windows that are
                         // not scored)
  const uint16_t* pos;   // qlen - qpos of the query's k-mers, grouped by code
  int32_t* h;            // this wavefront's KM_TGT_BINS bins
  int32_t ql, w, alpha, ncodes;
  uint32_t wm;           // floor((2^32 - 1) / w): n / w = umulhi(n, wm) or one more
  float rcp, hrcp;       // fl32(1 / w) and half of it
  uint32_t m24;          // ceil(2^25 / w): FDIV bins n / w = mul_hi_u24(128 n, m24)
  uint32_t w3;           // K = 3 dot-product weight bytes (DOT kernels)
  const uint8_t* flat;
  int64_t flat_len;
};

// Maximum of a 64-bit key over the wavefront (every lane active), in every
// lane: butterflies within each row of 16 lanes by DPP (quad_perm swaps,
// then row rotations by 4 and 8), then the four row maxima through
// readlane.  No LDS traffic.
__device__ __forceinline__ uint64_t km_wave_max_u64(uint64_t key) {
  auto step = [&](auto ctrl) {
    constexpr int C = decltype(ctrl)::value;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)key, C, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(key >> 32), C, 0xF, 0xF, false);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    key = o > key ? o : key;
  };
  step(std::integral_constant<int, 0xB1>{});   // quad_perm [1,0,3,2]
  step(std::integral_constant<int, 0x4E>{});   // quad_perm [2,3,0,1]
  step(std::integral_constant<int, 0x124>{});  // row_ror:4
  step(std::integral_constant<int, 0x128>{});  // row_ror:8
  uint64_t best = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, 16 * r);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(key >> 32), 16 * r);
    const uint64_t v = ((uint64_t)hi << 32) | lo;
    best = v > best ? v : best;
  }
  return best;
}

// One hit of target window p against query position entry qp (qlen - qpos):
// diagonal n = tpos - qpos + qlen >= 0 is counted in bin n / w, relative
// to the pass start b0 (ONEPASS: a single pass from bin 0 holds every bin).
// FDIV: n arrives times 128 (table entries and window positions carry the
// factor) and n / w = v_mul_hi_u32_u24(128 n, ceil(2^25 / w)): the high 32
// bits of the 48-bit product of the operands' low 24 bits, exact for n <
// 2^17 and 3 <= w <= 256 (checked for every such n and w); one instruction
// after the add that forms 128 n.  Otherwise a multiply-high and one
// correction.
template <bool FDIV, bool ONEPASS>
__device__ __forceinline__ void km_tgt_hit(const KmTgtQuery& Q, int32_t n, int32_t b0,
                                           int32_t nbp, bool on) {
  int32_t bin;
  if (FDIV) {
    asm("v_mul_hi_u32_u24 %0, %1, %2" : "=v"(bin) : "s"(Q.m24), "v"(n));
  } else {
    const uint32_t un = (uint32_t)n;
    uint32_t dq = __umulhi(un, Q.wm);
    dq += (un - dq * (uint32_t)Q.w >= (uint32_t)Q.w) ? 1u : 0u;
    bin = (int32_t)dq;
  }
  if (!ONEPASS) {
    bin -= b0;
    on = on && (uint32_t)bin < (uint32_t)nbp;
  }
  if (on) atomicAdd(&Q.h[bin], 1);
}

// Hits of the windows p, p+1, p+2, p+3 whose k-mers occur more than once in
// the query (entries with bit 31): two occurrences come in the entry itself
// (no loop, no read of pos), longer runs from pos.
template <bool FDIV, bool ONEPASS>
__device__ __forceinline__ void km_multi_hits(const KmTgtQuery& Q, const uint32_t (&v)[4],
                                              int32_t p, int32_t b0, int32_t nbp) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t e = v[i];
    if ((int32_t)e >= 0) continue;
    if ((e & 0x8000u) == 0) {
      km_tgt_hit<FDIV, ONEPASS>(Q, (p + i + (int32_t)(e & 0x7FFFu)) * KM_NS<FDIV>, b0, nbp, true);
      km_tgt_hit<FDIV, ONEPASS>(Q, (p + i + (int32_t)((e >> 16) & 0x7FFFu)) * KM_NS<FDIV>, b0,
                                nbp, true);
    } else {
      const uint32_t end = (e >> 16) & 0x7FFFu;
      for (uint32_t x = e & 0x7FFFu; x < end; ++x)
        km_tgt_hit<FDIV, ONEPASS>(Q, (p + i + (int32_t)Q.pos[x]) * KM_NS<FDIV>, b0, nbp, true);
    }
  }
}

__device__ __forceinline__ uint32_t km_byte(const uint32_t (&d)[3], int j) {
  return (d[j >> 2] >> (8 * (j & 3))) & 0xFFu;
}

// Scores the windows of one residue chunk held in d (relative bytes
// [8 lane, 8 lane + 12) past the survivor's 8-aligned base), bins
// [b0, b0 + nbp) of the pass.  Four windows at a time: their code-table
// entries, then the bin updates of k-mers at one query position (the entry
// is the position); the positions of a k-mer that occurs more than once in
// the query follow in a loop over its run (rare).  A window that is not scored (outside the
// survivor, or holding a residue >= alpha) reads the empty entry just before
// the table: its code is or-ed with all ones.
template <int K, bool FDIV, bool DOT, bool ONEPASS>
__device__ __forceinline__ void km_tgt_chunk(const KmTgtQuery& Q, const uint32_t (&d)[3],
                                             int32_t p_first, int32_t ntw, int32_t b0,
                                             int32_t nbp) {
  // windows u = 0..7 start at byte u of d; window u is scored when
  // 0 <= p_first + u < ntw and none of its K residues is >= alpha
  const int32_t lo = max(0, -p_first), hi = min(8, ntw - p_first);
  if (hi <= lo) return;
  uint32_t vmask = ((1u << hi) - 1u) & ~((1u << lo) - 1u);
  // residues >= alpha (SWAR: byte & 0x7F plus 0x80 - alpha sets the byte's
  // top bit exactly when it is >= alpha, alpha <= 128); the per-window mask
  // is built only when this lane holds such a residue
  const uint32_t add = (0x80u - (uint32_t)Q.alpha) * 0x01010101u;
  uint32_t t[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) t[x] = ((((d[x] & 0x7F7F7F7Fu) + add) | d[x]) & 0x80808080u) >> 7;
  if (__builtin_expect((t[0] | t[1] | t[2]) != 0, 0)) {
    uint32_t bm = 0;
#pragma unroll
    for (int x = 0; x < 3; ++x) bm |= ((t[x] * 0x00204081u) >> 21 & 0xFu) << (4 * x);
    uint32_t sb = bm;
#pragma unroll
    for (int j = 1; j < K; ++j) sb |= bm >> j;
    vmask &= ~sb;
    if (!vmask) return;
  }
  const uint32_t inv = ~vmask;
  // k-mer codes, exact for every byte value.  K = 3 with alpha <= 22: one
  // byte permute and one 4 x 8-bit dot product per window, bytes (r0, r0,
  // r1, r2) against weights (Q.w3: min(alpha^2, 255), the rest of alpha^2,
  // alpha, 1).  Otherwise Horner in base alpha (K <= 3) or the rolling form
  // c(u+1) = c(u) alpha - r(u) alpha^K + r(u+K); intermediate values stay
  // below 2^32 for alpha^K <= 8192.
  uint32_t cs[8];
  const uint32_t al = (uint32_t)Q.alpha & 0xFFu;   // alpha <= 128
  if constexpr (DOT) {
    static_assert(K == 3, "dot-product codes are built for K = 3");
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int l = u >> 2, o = u & 3;
      const uint32_t sel = (uint32_t)o | ((uint32_t)o << 8) | ((uint32_t)(o + 1) << 16) |
                           ((uint32_t)(o + 2) << 24);
      const uint32_t b = __builtin_amdgcn_perm(d[l + 1], d[l], sel);
      cs[u] = __builtin_amdgcn_udot4(b, Q.w3, 0u, false);
    }
  } else if constexpr (K <= 3) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      uint32_t c = km_byte(d, u);
#pragma unroll
      for (int j = 1; j < K; ++j) c = c * al + km_byte(d, u + j);
      cs[u] = c;
    }
  } else {
    uint32_t aK = 1;
#pragma unroll
    for (int j = 0; j < K; ++j) aK *= al;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) c = c * al + km_byte(d, j);
    cs[0] = c;
#pragma unroll
    for (int u = 1; u < 8; ++u) {
      c = c * al - km_byte(d, u - 1) * aK + km_byte(d, u + K - 1);
      cs[u] = c;
    }
  }
#pragma unroll
  for (int h = 0; h < 8; h += 4) {
    uint32_t v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t m = (uint32_t)((int32_t)(inv << (31 - (h + i))) >> 31);   // 0 or ~0
      v[i] = Q.tab[(int32_t)(cs[h + i] | m)];
    }
    // a k-mer at one query position: the entry is qlen - qpos itself
#pragma unroll
    for (int i = 0; i < 4; ++i)
      km_tgt_hit<FDIV, ONEPASS>(Q, p_first * KM_NS<FDIV> + (int32_t)v[i] + (h + i) * KM_NS<FDIV>, b0, nbp,
                                (int32_t)v[i] > 0);
    // bit 31: the k-mer occurs more than once in the query
    if (__builtin_expect(((v[0] | v[1] | v[2] | v[3]) >> 31) != 0, 0))
      km_multi_hits<FDIV, ONEPASS>(Q, v, p_first + h, b0, nbp);
  }
}

// One survivor: its first chunk is already in d0.  Returns the wavefront's
// key count << 32 | (2^32 - 1 - bin) of the best bin (all lanes).  The
// wavefront's bins are zero on entry and on return: the scan that reads a
// pass's bins clears them.
template <int K, bool FDIV, bool DOT>
__device__ __forceinline__ uint64_t km_tgt_survivor(const KmTgtQuery& Q, const uint32_t (&d0)[3],
                                                    int32_t tl, int64_t base, int32_t sh,
                                                    int lane) {
  const int32_t ntw = tl - K + 1;
  const int32_t nb = (FDIV ? (int32_t)__builtin_fmaf((float)(tl - K + Q.ql), Q.rcp, Q.hrcp)
                           : (tl - K + Q.ql) / Q.w) + 1;
  uint64_t key = 0;
  for (int32_t b0 = 0; b0 < nb; b0 += KM_TGT_BINS) {
    const int32_t nbp = min(KM_TGT_BINS, nb - b0);
    const bool one = nb <= KM_TGT_BINS;
    if (one) km_tgt_chunk<K, FDIV, DOT, true>(Q, d0, 8 * lane - sh, ntw, 0, nbp);
    else km_tgt_chunk<K, FDIV, DOT, false>(Q, d0, 8 * lane - sh, ntw, b0, nbp);
    for (int32_t r0 = KM_TGT_CHUNK; r0 < sh + ntw; r0 += KM_TGT_CHUNK) {
      const int32_t r = r0 + 8 * lane;
      uint32_t d[3] = {0, 0, 0};
      if (r + 7 - sh >= 0 && r - sh < ntw) km_load12(Q.flat, Q.flat_len, base + r, d);
      if (one) km_tgt_chunk<K, FDIV, DOT, true>(Q, d, r - sh, ntw, 0, nbp);
      else km_tgt_chunk<K, FDIV, DOT, false>(Q, d, r - sh, ntw, b0, nbp);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int32_t x = lane; x < nbp; x += WAVE) {
      const uint64_t kc = ((uint64_t)(uint32_t)Q.h[x] << 32) |
                          (uint32_t)(0xFFFFFFFFu - (uint32_t)(b0 + x));
      Q.h[x] = 0;
      key = kc > key ? kc : key;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  return km_wave_max_u64(key);
}

// Word W (0: low, 1: high) of d, zero-extended, times 4: one v_lshlrev_b32
// with an SDWA operand selection.
template <int W>
__device__ __forceinline__ uint32_t km_word_x4(uint32_t d) {
  uint32_t r;
  if (W == 0)
    asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
        "src1_sel:WORD_0"
        : "=v"(r)
        : "v"(d));
  else
    asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD "
        "src1_sel:WORD_1"
        : "=v"(r)
        : "v"(d));
  return r;
}

// Precomputed window codes (the database's 16-bit code array: code + 1, 0
// for windows that are not scored): NWIN windows of this lane from NWIN / 2
// dwords, starting at window p_first of the survivor.  Same hits and bins as
// km_tgt_chunk, without the code arithmetic: a stored code indexes the table
// from its empty entry (tab[-1], LDS address 0), so the entry's byte address
// is the stored code times 4.
template <int NWIN, bool FDIV, bool ONEPASS>
__device__ __forceinline__ void km_cd_chunk(const KmTgtQuery& Q, const uint32_t (&d)[NWIN / 2],
                                            int32_t p_first, int32_t b0, int32_t nbp) {
#pragma unroll
  for (int h = 0; h < NWIN; h += 4) {
    // LDS addresses of the entries: each 16-bit code times 4 (one SDWA
    // shift), the table starting at LDS address 0
    const uint32_t c[4] = {km_word_x4<0>(d[h / 2]), km_word_x4<1>(d[h / 2]),
                           km_word_x4<0>(d[h / 2 + 1]), km_word_x4<1>(d[h / 2 + 1])};
    uint32_t v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = *(const __attribute__((address_space(3))) uint32_t*)(size_t)c[i];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      km_tgt_hit<FDIV, ONEPASS>(Q, p_first * KM_NS<FDIV> + (int32_t)v[i] + (h + i) * KM_NS<FDIV>, b0, nbp,
                                (int32_t)v[i] > 0);
    if (__builtin_expect(((v[0] | v[1] | v[2] | v[3]) >> 31) != 0, 0))
      km_multi_hits<FDIV, ONEPASS>(Q, v, p_first + h, b0, nbp);
  }
}

// The windows of one survivor from its codes cs (npad codes, a multiple of
// 8, the last ones -1) into bins [b0, b0 + nbp); d0 holds this lane's codes
// 8 lane .. 8 lane + 7 (-1 past npad).  Full chunks give each lane 8
// windows (512 per wavefront); a last part of at most 256 windows gives each
// lane 4, and a survivor of at most 256 windows takes its codes from d0
// (lanes 32-63 the upper half of lanes 0-31).
template <bool FDIV, bool ONEPASS>
__device__ __forceinline__ void km_cd_pass(const KmTgtQuery& Q, const int16_t* cs, int32_t npad,
                                           const uint32_t (&d0)[4], int lane, int32_t b0,
                                           int32_t nbp) {
  if (npad <= 4 * WAVE) {
    const int src = lane & 31;
    const uint32_t z0 = __shfl(d0[2], src), z1 = __shfl(d0[3], src);
    const uint32_t h2[2] = {lane < 32 ? d0[0] : z0, lane < 32 ? d0[1] : z1};
    km_cd_chunk<4, FDIV, ONEPASS>(Q, h2, 8 * src + (lane < 32 ? 0 : 4), b0, nbp);
    return;
  }
  km_cd_chunk<8, FDIV, ONEPASS>(Q, d0, 8 * lane, b0, nbp);
  int32_t r0 = 8 * WAVE;
  for (; npad - r0 > 4 * WAVE; r0 += 8 * WAVE) {
    const int32_t p = r0 + 8 * lane;
    uint32_t d[4] = {0u, 0u, 0u, 0u};
    if (p < npad) {
      const uint4 v = *reinterpret_cast<const uint4*>(cs + p);
      d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
    }
    km_cd_chunk<8, FDIV, ONEPASS>(Q, d, p, b0, nbp);
  }
  if (r0 < npad) {
    const int32_t p = r0 + 4 * lane;
    uint32_t d[2] = {0u, 0u};
    if (p < npad) {
      const uint2 v = *reinterpret_cast<const uint2*>(cs + p);
      d[0] = v.x; d[1] = v.y;
    }
    km_cd_chunk<4, FDIV, ONEPASS>(Q, d, p, b0, nbp);
  }
}

// Reads and clears bins [0, nbp) of the wavefront's histogram (pass start
// b0) into the running key: count << 32 | (2^32 - 1 - bin).
__device__ __forceinline__ uint64_t km_tgt_scan(const KmTgtQuery& Q, int32_t b0, int32_t nbp,
                                                int lane, uint64_t key) {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int32_t x = lane; x < nbp; x += WAVE) {
    const uint64_t kc = ((uint64_t)(uint32_t)Q.h[x] << 32) |
                        (uint32_t)(0xFFFFFFFFu - (uint32_t)(b0 + x));
    Q.h[x] = 0;
    key = kc > key ? kc : key;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return key;
}

// One survivor from its window codes; returns the best bin's key as
// km_tgt_survivor does.  Survivors whose bins fit one histogram pass (nearly
// all) take a path without the per-hit pass test.
template <int K, bool FDIV>
__device__ __forceinline__ uint64_t km_cd_survivor(const KmTgtQuery& Q, const int16_t* cs,
                                                   int32_t npad, const uint32_t (&d0)[4],
                                                   int32_t tl, int lane) {
  const int32_t nb = (FDIV ? (int32_t)__builtin_fmaf((float)(tl - K + Q.ql), Q.rcp, Q.hrcp)
                           : (tl - K + Q.ql) / Q.w) + 1;
  uint64_t key = 0;
  if (nb <= KM_TGT_BINS) {
    km_cd_pass<FDIV, true>(Q, cs, npad, d0, lane, 0, nb);
    key = km_tgt_scan(Q, 0, nb, lane, key);
  } else {
    for (int32_t b0 = 0; b0 < nb; b0 += KM_TGT_BINS) {
      const int32_t nbp = min(KM_TGT_BINS, nb - b0);
      km_cd_pass<FDIV, false>(Q, cs, npad, d0, lane, b0, nbp);
      key = km_tgt_scan(Q, b0, nbp, lane, key);
    }
  }
  return km_wave_max_u64(key);
}

// CODES: the survivors' windows come from the database's precomputed window
// codes (codes, dboff = code offsets, multiples of 8) instead of their
// residues (dbflat, dboff = residue offsets).
template <int K, bool FDIV, bool DOT, bool CODES>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8))) void km_phase_b_target_kernel(
    const int32_t* __restrict__ qcode, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, int32_t alpha, int32_t ncodes,
    const uint8_t* __restrict__ dbflat, int64_t dbflat_len, const int16_t* __restrict__ codes,
    const int64_t* __restrict__ dboff,
    const int32_t* __restrict__ dblen, const int32_t* __restrict__ surv,
    const int32_t* __restrict__ surv_n, int32_t Ks, int32_t w, int32_t per_wg, int32_t max_nk,
    uint32_t* __restrict__ best, int32_t* __restrict__ best_bin) {
  // Dynamic LDS only, so the code table starts at LDS address 0 (the host
  // checks that the kernel has no static LDS) and a window's table entry
  // sits at LDS address 4 x its stored code: table, query positions, the
  // wavefronts' bins, then the survivor metadata.
  extern __shared__ int32_t km_lds[];
  int32_t* const tail = km_lds + ((ncodes + 2 + (max_nk + 1) / 2 + 16 * KM_TGT_BINS) & ~1);
  int64_t* const m_off = reinterpret_cast<int64_t*>(tail);   // 8-byte aligned
  int32_t* const m_tl = tail + 2 * KM_TGT_PER_WG;
  int32_t* const m_pre = m_tl + KM_TGT_PER_WG;
  int32_t* const ws = m_pre + KM_TGT_PER_WG;
  int32_t& m_total = ws[32];
  const int q = blockIdx.y;
  const int32_t ns = surv_n[q];
  const int32_t s_lo = (int32_t)blockIdx.x * per_wg, s_hi = min(ns, s_lo + per_wg);
  if (s_lo >= s_hi) return;
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  // survivor metadata of the workgroup's list (at most KM_TGT_PER_WG <=
  // blockDim.x entries, one per thread) and the exclusive prefix of their
  // costs
  {
    int32_t cost = 0;
    if ((int32_t)threadIdx.x < s_hi - s_lo) {
      const int32_t t = surv[(int64_t)q * Ks + s_lo + threadIdx.x];
      const int32_t len = dblen[t];
      m_tl[threadIdx.x] = len;
      m_off[threadIdx.x] = dboff[t];
      cost = len + 64;
    }
    int32_t total;
    const int32_t pre = block_excl_scan(cost, ws, &total);
    if ((int32_t)threadIdx.x < s_hi - s_lo) m_pre[threadIdx.x] = pre;
    if (threadIdx.x == 0) m_total = total;
  }
  int32_t* tab = km_lds + 1;               // [ncodes]; km_lds[0] = tab[-1] stays 0
  uint16_t* pos = reinterpret_cast<uint16_t*>(km_lds + ncodes + 1);   // [max_nk]
  const int32_t ql = qlen[q];
  const int32_t nk = min(ql - K + 1, max_nk);   // max_nk bounds every query of the batch
  const int32_t* qc = qcode + qoff[q];

  // counting sort of the query's usable k-mer positions by code
  for (int32_t c = (int32_t)threadIdx.x - 1; c < ncodes; c += blockDim.x) tab[c] = 0;
  __syncthreads();
  for (int32_t p = threadIdx.x; p < nk; p += blockDim.x) {
    const int32_t c = qc[p];
    if (c >= 0) atomicAdd(&tab[c], 1);
  }
  __syncthreads();
  // KM_TGT_CODES / 1024 = 8 codes per thread at most
  const int32_t per = (ncodes + (int32_t)blockDim.x - 1) / (int32_t)blockDim.x;
  const int32_t c0 = (int32_t)threadIdx.x * per, c1 = min(ncodes, c0 + per);
  int32_t loc = 0;
  for (int32_t c = c0; c < c1; ++c) loc += tab[c];
  int32_t tot;
  const int32_t first = block_excl_scan(loc, ws, &tot);
  int32_t run = first;
  for (int32_t c = c0; c < c1; ++c) {
    const int32_t v = tab[c];
    tab[c] = run;
    run += v;
  }
  __syncthreads();
  for (int32_t p = threadIdx.x; p < nk; p += blockDim.x) {
    const int32_t c = qc[p];
    if (c >= 0) pos[atomicAdd(&tab[c], 1)] = (uint16_t)(ql - p);
  }
  __syncthreads();
  // tab[c] now holds the end of c's run; its start is the previous end
  // (first for the thread's first code).  Entry of code c: 0 when the query
  // lacks it; its qlen - qpos (1 .. 16385) when it occurs once; bit 31 | a |
  // b << 16, its two qlen - qpos values, when it occurs twice; else bit 31 |
  // bit 15 | start | end << 16, its run in pos (element indices, nk <=
  // 16383).  tab[-1] stays 0 (empty).
  {
    int32_t prev = first;
    for (int32_t c = c0; c < c1; ++c) {
      const int32_t e = tab[c];
      const int32_t n = e - prev;
      tab[c] = n == 0   ? 0
               : n == 1 ? (int32_t)pos[prev] * KM_NS<FDIV>
               : n == 2 ? (int32_t)(0x80000000u | (uint32_t)pos[prev] |
                                    ((uint32_t)pos[prev + 1] << 16))
                        : (int32_t)(0x80008000u | (uint32_t)prev | ((uint32_t)e << 16));
      prev = e;
    }
  }
  __syncthreads();

  KmTgtQuery Q;
  Q.tab = reinterpret_cast<const uint32_t*>(tab);
  Q.pos = pos;
  Q.h = km_lds + ncodes + 1 + (max_nk + 1) / 2 + wv * KM_TGT_BINS;
  Q.ql = ql;
  Q.w = w;
  Q.wm = 0xFFFFFFFFu / (uint32_t)w;
  Q.rcp = (float)(1.0 / (double)w);
  Q.hrcp = 0.5f * Q.rcp;
  Q.m24 = ((1u << 25) + (uint32_t)w - 1u) / (uint32_t)w;
  Q.alpha = alpha;
  Q.ncodes = ncodes;
  {
    // DOT is launched for K = 3 and alpha^2 <= 510 only
    const uint32_t a2 = (uint32_t)(alpha * alpha);
    Q.w3 = DOT ? (min(a2, 255u) | ((a2 - min(a2, 255u)) << 8) | ((uint32_t)alpha << 16) |
                  (1u << 24))
               : 0u;
  }
  Q.flat = dbflat;
  Q.flat_len = dbflat_len;
  for (int32_t x = lane; x < KM_TGT_BINS; x += WAVE) Q.h[x] = 0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int32_t cnt = s_hi - s_lo;
  // cost-balanced contiguous ranges: wavefront wv takes the survivors whose
  // cost prefix (length + 64 each) falls in [wv T / nw, (wv + 1) T / nw)
  int32_t lb[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int64_t t = (int64_t)(wv + e) * m_total / (blockDim.x / WAVE);
    int32_t a = 0, b = cnt;
    while (a < b) {
      const int32_t mid = (a + b) >> 1;
      if ((int64_t)m_pre[mid] < t) a = mid + 1;
      else b = mid;
    }
    lb[e] = __builtin_amdgcn_readfirstlane(a);   // equal in every lane already
  }
  constexpr int G = CODES ? 3 : 4;
  for (int32_t i0 = lb[0]; i0 < lb[1]; i0 += G) {
    // first chunks of G survivors in flight together
    constexpr int DW = CODES ? 4 : 3;   // dwords of a lane's first chunk
    uint32_t d[G][DW];
    int32_t tl[G], sh[G];
    int64_t base[G];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      const int32_t i = min(i0 + gi, lb[1] - 1);
      tl[gi] = __builtin_amdgcn_readfirstlane(m_tl[i]);
      const int64_t off0 = m_off[i];
      const int64_t off = ((int64_t)__builtin_amdgcn_readfirstlane((int32_t)(off0 >> 32)) << 32) |
                          (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)off0);
      if constexpr (CODES) {
        // sh: the survivor's padded code count
        sh[gi] = (max(tl[gi] - K + 1, 0) + 7) & ~7;
        base[gi] = off;
#pragma unroll
        for (int x = 0; x < DW; ++x) d[gi][x] = 0u;
        if (i0 + gi < lb[1] && 8 * lane < sh[gi]) {
          const uint4 v = *reinterpret_cast<const uint4*>(codes + off + 8 * lane);
          d[gi][0] = v.x; d[gi][1] = v.y; d[gi][2] = v.z; d[gi][DW - 1] = v.w;
        }
      } else {
        sh[gi] = (int32_t)(reinterpret_cast<uintptr_t>(dbflat + off) & 7);
        base[gi] = off - sh[gi];
#pragma unroll
        for (int x = 0; x < DW; ++x) d[gi][x] = 0;
        const int32_t r = 8 * lane;
        if (i0 + gi < lb[1] && r + 7 - sh[gi] >= 0 && r - sh[gi] < tl[gi] - K + 1) {
          uint32_t d3[3];
          km_load12(dbflat, dbflat_len, base[gi] + r, d3);
#pragma unroll
          for (int x = 0; x < DW; ++x) d[gi][x] = d3[x];
        }
      }
    }
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      if (i0 + gi >= lb[1]) break;
      uint64_t key;
      if constexpr (CODES) {
        const uint32_t d4[4] = {d[gi][0], d[gi][1], d[gi][2], d[gi][DW - 1]};
        key = km_cd_survivor<K, FDIV>(Q, codes + base[gi], sh[gi], d4, tl[gi], lane);
      } else {
        const uint32_t d3[3] = {d[gi][0], d[gi][1], d[gi][DW - 1]};
        key = km_tgt_survivor<K, FDIV, DOT>(Q, d3, tl[gi], base[gi], sh[gi], lane);
      }
      if (lane == 0) {
        const int64_t o = (int64_t)q * Ks + s_lo + i0 + gi;
        best[o] = (uint32_t)(key >> 32);
        best_bin[o] = key == 0 ? 0 : (int32_t)(0xFFFFFFFFu - (uint32_t)key);
      }
    }
  }
}
#define KM_TGT_INST(K, F, D, C)                                                               \
  template __global__ void km_phase_b_target_kernel<K, F, D, C>(                              \
      const int32_t*, const int64_t*, const int32_t*, int32_t, int32_t, const uint8_t*,       \
      int64_t, const int16_t*, const int64_t*, const int32_t*, const int32_t*, const int32_t*, \
      int32_t, int32_t, int32_t, int32_t, uint32_t*, int32_t*);
KM_TGT_INST(1, false, false, false)
KM_TGT_INST(2, false, false, false)
KM_TGT_INST(3, false, false, false)
KM_TGT_INST(4, false, false, false)
KM_TGT_INST(5, false, false, false)
KM_TGT_INST(1, true, false, false)
KM_TGT_INST(2, true, false, false)
KM_TGT_INST(3, true, false, false)
KM_TGT_INST(4, true, false, false)
KM_TGT_INST(5, true, false, false)
KM_TGT_INST(3, false, true, false)
KM_TGT_INST(3, true, true, false)
KM_TGT_INST(1, false, false, true)
KM_TGT_INST(2, false, false, true)
KM_TGT_INST(3, false, false, true)
KM_TGT_INST(4, false, false, true)
KM_TGT_INST(5, false, false, true)
KM_TGT_INST(1, true, false, true)
KM_TGT_INST(2, true, false, true)
KM_TGT_INST(3, true, false, true)
KM_TGT_INST(4, true, false, true)
KM_TGT_INST(5, true, false, true)
#undef KM_TGT_INST

// Final ordering: when more than max_cands were kept, sort the selected
// entries by (score desc, id asc) (bitonic sort in LDS, one workgroup per
// query); otherwise keep ascending id.  Writes targets/scores/diags.
// src_ids: selected positions (ascending) into the dense score array;
// map_ids: optional position -> target id map (survivor list);
// bins: optional best bin per position (diag hint = bin*w - qlen + w/2).
__global__ __launch_bounds__(1024) void km_finalize_kernel(
    const int32_t* __restrict__ sel_ids, const int32_t* __restrict__ sel_n,
    const int32_t* __restrict__ sel_trunc, int32_t K, const uint32_t* __restrict__ scores,
    int64_t score_stride, const int32_t* __restrict__ map_ids, int32_t map_stride,
    const int32_t* __restrict__ bins, const int32_t* __restrict__ qlen, int32_t w,
    int32_t max_cands, int32_t* __restrict__ out_targets, int32_t* __restrict__ out_counts,
    int32_t* __restrict__ out_scores, int32_t* __restrict__ out_diags) {
  extern __shared__ uint64_t keys[];   // [pow2 >= K]
  const int q = blockIdx.x, tid = threadIdx.x;
  const int32_t n = sel_n[q];
  const int32_t* si = sel_ids + (int64_t)q * K;
  const uint32_t* sc = scores + (int64_t)q * score_stride;
  int32_t npow = 1;
  while (npow < n) npow <<= 1;
  const bool sort = sel_trunc[q] != 0;
  for (int x = tid; x < npow; x += blockDim.x) {
    uint64_t key = ~0ull;   // padding sorts last
    if (x < n) {
      const int32_t pos = si[x];
      const uint32_t v = min(sc[pos] & 0x7FFFFFFFu, (uint32_t)KM_SAT);
      // ascending key order == (score desc, position asc)
      key = ((uint64_t)(KM_SAT - v) << 32) | (uint32_t)pos;
    }
    keys[x] = key;
  }
  __syncthreads();
  if (sort) {
    for (int32_t size = 2; size <= npow; size <<= 1) {
      for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (int x = tid; x < npow; x += blockDim.x) {
          const int32_t y = x ^ stride;
          if (y > x) {
            const bool up = (x & size) == 0;
            const uint64_t a = keys[x], b = keys[y];
            if ((a > b) == up) { keys[x] = b; keys[y] = a; }
          }
        }
        __syncthreads();
      }
    }
  }
  const int32_t ql = qlen[q];
  for (int x = tid; x < n; x += blockDim.x) {
    const uint64_t key = keys[x];
    const int32_t pos = (int32_t)(key & 0xFFFFFFFFu);
    const int32_t v = KM_SAT - (int32_t)(key >> 32);
    const int64_t o = (int64_t)q * max_cands + x;
    out_targets[o] = map_ids ? map_ids[(int64_t)q * map_stride + pos] : pos;
    out_scores[o] = v;
    out_diags[o] = bins ? bins[(int64_t)q * map_stride + pos] * w - ql + w / 2 : 0;
  }
  if (tid == 0) out_counts[q] = n;
}

}  // namespace ck
