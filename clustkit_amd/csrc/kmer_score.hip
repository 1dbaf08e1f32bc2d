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

// Phase A, tiled: grid (tiles, queries).  counts[q*nd + t] = raw hit count.
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
  for (int32_t p = 0; p < nk; ++p) {
    const int32_t c = qc[p];
    if (c < 0) continue;
    const int64_t* to = tile_off + (int64_t)c * (ntiles + 1) + j;
    const int64_t e = to[1];
    for (int64_t h = to[0] + threadIdx.x; h < e; h += blockDim.x)
      atomicAdd(&cnt[ptid[h] - t0], 1u);
  }
  __syncthreads();
  const int32_t n = min(KM_TILE, nd - t0);
  uint32_t* out = counts + (int64_t)q * nd + t0;
  for (int x = threadIdx.x; x < n; x += blockDim.x) out[x] = cnt[x];
}

// Phase A, global atomics: grid (chunks, queries) over the query's k-mers.
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
// ascending id order.  One workgroup of 1024 threads per query.  Scores are
// read as min(score & 0x7FFFFFFF, 32767); only ids with score >= min_score
// take part.  Optionally marks the selected ids (bit 31 of the score) and
// emits per-selected prefix offsets of nbins(id) (Phase-B histogram layout).
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
  int32_t mark;           // set bit 31 on selected entries
  // Phase-B histogram offsets (optional): nbins(id) = (len[id] - k + qlen)/w + 1
  int32_t* out_hoff;      // [nq][K+1] or nullptr
  const int32_t* id_len;  // target lengths
  const int32_t* qlen;
  int32_t k, w;
};

__global__ __launch_bounds__(1024) void km_select_kernel(SelectArgs a) {
  __shared__ int32_t hist[KM_SAT + 1];
  __shared__ int32_t wsum[3][16];
  __shared__ int32_t s_cut, s_take, s_total;
  const int q = blockIdx.x;
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  uint32_t* sc = a.scores + (int64_t)q * a.stride;
  int32_t n = a.n_of_q ? a.n_of_q[q] : a.n_fixed;
  if (a.qlen && a.qlen[q] < a.k) n = 0;   // no k-mer in the query: no candidate
  for (int x = tid; x <= KM_SAT; x += blockDim.x) hist[x] = 0;
  __syncthreads();
  for (int32_t t = tid; t < n; t += blockDim.x) {
    const int32_t v = (int32_t)min(sc[t] & 0x7FFFFFFFu, (uint32_t)KM_SAT);
    if (v >= a.min_score) atomicAdd(&hist[v], 1);
  }
  __syncthreads();
  // Suffix sums over the histogram: each thread owns 32 consecutive bins.
  if (tid == 0) { s_total = 0; s_cut = -1; s_take = 0; }
  __syncthreads();
  {
    int32_t local = 0;
    for (int b = 0; b < 32; ++b) local += hist[tid * 32 + b];
    // inclusive suffix over threads: sum of local for threads >= tid
    __shared__ int32_t part[1024];
    part[tid] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int32_t add = tid + off < 1024 ? part[tid + off] : 0;
      __syncthreads();
      part[tid] += add;
      __syncthreads();
    }
    if (tid == 0) s_total = part[0];
    __syncthreads();
    const int32_t total = s_total;
    if (total > a.K) {
      // bins above this thread's block hold `above`; find the cut inside
      int32_t above = part[tid] - local;
      if (above < a.K && part[tid] >= a.K) {
        for (int b = 31; b >= 0; --b) {
          const int32_t h = hist[tid * 32 + b];
          if (above + h >= a.K) { s_cut = tid * 32 + b; s_take = a.K - above; break; }
          above += h;
        }
      }
    }
  }
  __syncthreads();
  const int32_t total = s_total;
  const bool trunc = total > a.K;
  const int32_t cut = trunc ? s_cut : a.min_score;   // scores > cut always taken
  const int32_t take = trunc ? s_take : 0;           // how many == cut (lowest ids)
  // Ordered emission in ascending id with a running base.
  int32_t emitted = 0, ties = 0, hbase = 0;
  int32_t* oid = a.out_ids + (int64_t)q * a.K;
  int32_t* ohoff = a.out_hoff ? a.out_hoff + (int64_t)q * (a.K + 1) : nullptr;
  const int32_t ql = a.qlen ? a.qlen[q] : 0;
  for (int32_t t0 = 0; t0 < n; t0 += blockDim.x) {
    const int32_t t = t0 + tid;
    int32_t v = -1;
    if (t < n) v = (int32_t)min(sc[t] & 0x7FFFFFFFu, (uint32_t)KM_SAT);
    const bool in = v >= a.min_score;
    const bool above = trunc ? (in && v > cut) : in;
    const bool tie = trunc && in && v == cut;
    const uint64_t ba = __ballot(above), bt = __ballot(tie);
    const uint64_t lm = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int32_t pa = __popcll(ba & lm), pt = __popcll(bt & lm);
    int32_t nb = 0;
    if (lane == 0) { wsum[0][wv] = __popcll(ba); wsum[1][wv] = __popcll(bt); }
    __syncthreads();
    int32_t wa = 0, wt = 0, tota = 0, tott = 0;
    for (int x = 0; x < 16; ++x) {
      if (x < wv) { wa += wsum[0][x]; wt += wsum[1][x]; }
      tota += wsum[0][x];
      tott += wsum[1][x];
    }
    const int32_t tie_rank = ties + wt + pt;
    const bool sel = above || (tie && tie_rank < take);
    // selected rank among this chunk: above-count before + selected ties before
    const int32_t ties_taken_before = max(0, min(take - ties, wt + pt));
    const int32_t rank = emitted + (wa + pa) + (trunc ? ties_taken_before : 0);
    // nbins prefix needs a second ordered scan: per-lane value if selected
    if (sel && ohoff) nb = (a.id_len[t] - a.k + ql) / a.w + 1;
    __syncthreads();
    // exclusive scan of nb over the block (only selected contribute)
    __shared__ int32_t nbs[1024];
    nbs[tid] = nb;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int32_t add = tid >= off ? nbs[tid - off] : 0;
      __syncthreads();
      nbs[tid] += add;
      __syncthreads();
    }
    if (sel) {
      oid[rank] = t;
      if (ohoff) ohoff[rank] = hbase + nbs[tid] - nb;
      if (a.mark) sc[t] |= KM_SURV;
    }
    const int32_t chunk_nb = nbs[1023];
    const int32_t ties_here = tott;
    const int32_t taken_here = trunc ? max(0, min(take - ties, ties_here)) : 0;
    emitted += tota + taken_here;
    ties += ties_here;
    hbase += chunk_nb;
    __syncthreads();
  }
  if (tid == 0) {
    const int32_t cnt = trunc ? a.K : total;
    a.out_n[q] = cnt;
    if (a.out_trunc) a.out_trunc[q] = trunc ? 1 : 0;
    if (ohoff) ohoff[cnt] = hbase;
  }
}

// Phase B, tiled: hits on survivor targets of the tile -> diagonal-bin
// histograms.  grid (tiles, queries).
__global__ __launch_bounds__(256) void km_phase_b_tiled_kernel(
    const int32_t* __restrict__ qcode, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, int32_t k, const int64_t* __restrict__ tile_off,
    int32_t ntiles, const int32_t* __restrict__ ptid, const int32_t* __restrict__ ppos,
    int32_t nd, const uint32_t* __restrict__ counts, const int32_t* __restrict__ surv_ids,
    const int32_t* __restrict__ surv_n, const int32_t* __restrict__ hoff, int32_t K,
    int32_t w, int32_t* __restrict__ hist, int64_t hist_stride) {
  __shared__ uint32_t flags[KM_TILE / 32];
  const int j = blockIdx.x, q = blockIdx.y;
  const int32_t t0 = j * KM_TILE;
  const int32_t n = min(KM_TILE, nd - t0);
  const uint32_t* cq = counts + (int64_t)q * nd + t0;
  for (int x = threadIdx.x; x < KM_TILE / 32; x += blockDim.x) flags[x] = 0;
  __syncthreads();
  bool any = false;
  for (int x = threadIdx.x; x < n; x += blockDim.x)
    if (cq[x] & KM_SURV) { atomicOr(&flags[x >> 5], 1u << (x & 31)); any = true; }
  if (!__syncthreads_or(any)) return;
  const int32_t ql = qlen[q];
  const int32_t nk = ql - k + 1;
  const int32_t* qc = qcode + qoff[q];
  const int32_t* sid = surv_ids + (int64_t)q * K;
  const int32_t ns = surv_n[q];
  const int32_t* ho = hoff + (int64_t)q * (K + 1);
  int32_t* hq = hist + (int64_t)q * hist_stride;
  for (int32_t p = 0; p < nk; ++p) {
    const int32_t c = qc[p];
    if (c < 0) continue;
    const int64_t* to = tile_off + (int64_t)c * (ntiles + 1) + j;
    const int64_t e = to[1];
    for (int64_t h = to[0] + threadIdx.x; h < e; h += blockDim.x) {
      const int32_t t = ptid[h];
      const int32_t x = t - t0;
      if (!(flags[x >> 5] & (1u << (x & 31)))) continue;
      int32_t lo = 0, hi = ns;                      // survivor slot of t
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (sid[mid] < t) lo = mid + 1; else hi = mid;
      }
      const int32_t bin = (ppos[h] - p + ql) / w;
      atomicAdd(&hq[ho[lo] + bin], 1);
    }
  }
}

// Phase B with global survivor flags (sparse indexes).
__global__ __launch_bounds__(256) void km_phase_b_global_kernel(
    const int32_t* __restrict__ qcode, const int64_t* __restrict__ qoff,
    const int32_t* __restrict__ qlen, int32_t k, const int64_t* __restrict__ koff,
    const int32_t* __restrict__ ptid, const int32_t* __restrict__ ppos, int32_t nd,
    const uint32_t* __restrict__ counts, const int32_t* __restrict__ surv_ids,
    const int32_t* __restrict__ surv_n, const int32_t* __restrict__ hoff, int32_t K,
    int32_t w, int32_t* __restrict__ hist, int64_t hist_stride) {
  const int q = blockIdx.y;
  const int32_t ql = qlen[q];
  const int32_t nk = ql - k + 1;
  const int32_t* qc = qcode + qoff[q];
  const uint32_t* cq = counts + (int64_t)q * nd;
  const int32_t* sid = surv_ids + (int64_t)q * K;
  const int32_t ns = surv_n[q];
  const int32_t* ho = hoff + (int64_t)q * (K + 1);
  int32_t* hq = hist + (int64_t)q * hist_stride;
  const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const int nw = blockDim.x / WAVE;
  for (int32_t p = blockIdx.x * nw + wv; p < nk; p += gridDim.x * nw) {
    const int32_t c = qc[p];
    if (c < 0) continue;
    for (int64_t h = koff[c] + lane; h < koff[c + 1]; h += WAVE) {
      const int32_t t = ptid[h];
      if (!(cq[t] & KM_SURV)) continue;
      int32_t lo = 0, hi = ns;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (sid[mid] < t) lo = mid + 1; else hi = mid;
      }
      atomicAdd(&hq[ho[lo] + (ppos[h] - p + ql) / w], 1);
    }
  }
}

// Best diagonal bin per survivor (first maximum); resets the bins it reads
// so the histogram pool is zero for the next batch.  grid (chunks, queries).
__global__ __launch_bounds__(256) void km_phase_b_best_kernel(
    const int32_t* __restrict__ surv_n, const int32_t* __restrict__ hoff, int32_t K,
    int32_t* __restrict__ hist, int64_t hist_stride, uint32_t* __restrict__ best,
    int32_t* __restrict__ best_bin) {
  const int q = blockIdx.y;
  const int32_t ns = surv_n[q];
  const int32_t* ho = hoff + (int64_t)q * (K + 1);
  int32_t* hq = hist + (int64_t)q * hist_stride;
  for (int32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
    int32_t bc = 0, bb = 0;
    for (int32_t x = ho[s]; x < ho[s + 1]; ++x) {
      const int32_t v = hq[x];
      if (v > bc) { bc = v; bb = x - ho[s]; }
      if (v) hq[x] = 0;
    }
    best[(int64_t)q * K + s] = (uint32_t)bc;
    best_bin[(int64_t)q * K + s] = bb;
  }
}

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
