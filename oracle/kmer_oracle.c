/*
 * CPU oracle for k-mer index scoring -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates the semantics of clustkit/csrc/kmer_score.c:198-472
 * (score_query_full / batch_score_queries_c) in its own formulation:
 *
 *   Phase A  count[t] = number of (query k-mer, posting) hits on target t,
 *            saturating at 32767; a query k-mer counts only when all its
 *            residues are < alpha and its frequency is <= freq_thresh.
 *   pass     count[t] >= min_total_hits, ascending target id.
 *   top-K    (count desc, id asc); when no more than K pass, ascending id.
 *   Phase B  for every surviving target, histogram of diagonal bins
 *            (tpos - qpos + qlen) / w over its hits; best bin = highest
 *            count, lowest bin on ties; kept when >= min_diag_hits; hint =
 *            bin * w - qlen + w / 2.
 *   output   top max_cands of the kept targets by (best run desc, id asc),
 *            ascending id when no more than max_cands are kept.
 * With min_diag_hits <= 1 Phase B is skipped and the Phase-A counts are
 * ranked directly (hints are written as 0; the reference leaves them
 * unwritten in that mode).
 *
 * The reference sorts packed (tid * 10^6 + bin) keys and counts runs; this
 * restatement histograms per target instead -- same result, different
 * mechanics, pinned against golden vectors from the reference object.
 *
 * kmer_similar_oracle_batch (end of file) restates the BLOSUM-similar 5-mer
 * stage 1 of kmer_index.py:1186-1330; it is pinned by an enumeration of all
 * 20^5 codes per query position and by a closed form (tests/test_kmer_oracle.py).
 */

#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef struct { int32_t id; int32_t score; int32_t diag; } korc_cand;

static int korc_cmp(const void* x, const void* y) {
    const korc_cand* a = (const korc_cand*)x;
    const korc_cand* b = (const korc_cand*)y;
    if (a->score != b->score) return a->score > b->score ? -1 : 1;
    return (a->id > b->id) - (a->id < b->id);
}

/* keep the first `k` of cands (n entries, ascending id): sort only when
 * truncation is needed, as the reference's topk_by_score does. */
static int32_t korc_top(korc_cand* c, int32_t n, int32_t k) {
    if (n <= k) return n;
    qsort(c, (size_t)n, sizeof(korc_cand), korc_cmp);
    return k;
}

static int32_t korc_query(const uint8_t* q, int32_t qlen, int32_t k, int32_t alpha,
                          const int64_t* koff, const int64_t* kent, const int32_t* kfreq,
                          int32_t fthr, int32_t ndb, int32_t min_total, int32_t min_diag,
                          int32_t w, int32_t topk_a, int32_t max_cands,
                          int32_t* o_ids, int32_t* o_sc, int32_t* o_dg) {
    int32_t nk = qlen - k + 1;
    if (nk <= 0) return 0;
    int64_t* code = (int64_t*)malloc(sizeof(int64_t) * (size_t)nk);
    for (int32_t p = 0; p < nk; p++) {
        int64_t c = 0;
        for (int32_t j = 0; j < k; j++) {
            uint8_t r = q[p + j];
            if (r >= alpha) { c = -1; break; }
            c = c * alpha + r;
        }
        if (c >= 0 && kfreq[c] > fthr) c = -1;
        code[p] = c;
    }
    int32_t* cnt = (int32_t*)calloc((size_t)ndb, sizeof(int32_t));
    for (int32_t p = 0; p < nk; p++) {
        if (code[p] < 0) continue;
        for (int64_t h = koff[code[p]]; h < koff[code[p] + 1]; h++) {
            int32_t t = (int32_t)(kent[h] >> 32);
            if (cnt[t] < 32767) cnt[t]++;
        }
    }
    int32_t npass = 0;
    for (int32_t t = 0; t < ndb; t++) npass += cnt[t] >= min_total;
    korc_cand* c = (korc_cand*)malloc(sizeof(korc_cand) * (size_t)(npass ? npass : 1));
    int32_t n = 0;
    for (int32_t t = 0; t < ndb; t++)
        if (cnt[t] >= min_total) { c[n].id = t; c[n].score = cnt[t]; c[n].diag = 0; n++; }
    int32_t out = 0;
    if (n == 0) {
        out = 0;
    } else if (min_diag <= 1) {
        out = korc_top(c, n, max_cands);
        for (int32_t i = 0; i < out; i++) { o_ids[i] = c[i].id; o_sc[i] = c[i].score; o_dg[i] = 0; }
    } else {
        int32_t ns = korc_top(c, n, topk_a);
        /* survivors in ascending id for the final ordering */
        int32_t* surv = (int32_t*)malloc(sizeof(int32_t) * (size_t)ns);
        for (int32_t i = 0; i < ns; i++) surv[i] = c[i].id;
        for (int32_t i = 1; i < ns; i++) {            /* insertion sort by id */
            int32_t v = surv[i], j = i - 1;
            while (j >= 0 && surv[j] > v) { surv[j + 1] = surv[j]; j--; }
            surv[j + 1] = v;
        }
        int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)ndb);
        for (int32_t t = 0; t < ndb; t++) slot[t] = -1;
        for (int32_t i = 0; i < ns; i++) slot[surv[i]] = i;
        /* per-survivor histogram over diagonal bins, sized by max bin */
        int32_t nbins = 0;
        for (int32_t p = 0; p < nk; p++) {
            if (code[p] < 0) continue;
            for (int64_t h = koff[code[p]]; h < koff[code[p] + 1]; h++) {
                int32_t t = (int32_t)(kent[h] >> 32);
                if (slot[t] < 0) continue;
                int32_t b = ((int32_t)(kent[h] & 0xFFFFFFFF) - p + qlen) / w;
                if (b + 1 > nbins) nbins = b + 1;
            }
        }
        int32_t* hist = (int32_t*)calloc((size_t)ns * (size_t)(nbins ? nbins : 1), sizeof(int32_t));
        for (int32_t p = 0; p < nk; p++) {
            if (code[p] < 0) continue;
            for (int64_t h = koff[code[p]]; h < koff[code[p] + 1]; h++) {
                int32_t t = (int32_t)(kent[h] >> 32);
                if (slot[t] < 0) continue;
                int32_t b = ((int32_t)(kent[h] & 0xFFFFFFFF) - p + qlen) / w;
                hist[(size_t)slot[t] * nbins + b]++;
            }
        }
        int32_t nf = 0;
        for (int32_t i = 0; i < ns; i++) {
            int32_t bestc = 0, bestb = 0;
            for (int32_t b = 0; b < nbins; b++) {
                int32_t v = hist[(size_t)i * nbins + b];
                if (v > bestc) { bestc = v; bestb = b; }
            }
            if (bestc > 0 && bestc >= min_diag) {
                c[nf].id = surv[i]; c[nf].score = bestc;
                c[nf].diag = bestb * w - qlen + w / 2;
                nf++;
            }
        }
        out = korc_top(c, nf, max_cands);
        for (int32_t i = 0; i < out; i++) { o_ids[i] = c[i].id; o_sc[i] = c[i].score; o_dg[i] = c[i].diag; }
        free(hist); free(slot); free(surv);
    }
    free(c); free(cnt); free(code);
    return out;
}

void kmer_oracle_batch(const uint8_t* qflat, const int64_t* qoff, const int32_t* qlen, int32_t nq,
                       int32_t k, int32_t alpha, const int64_t* koff, const int64_t* kent,
                       const int32_t* kfreq, int32_t fthr, int32_t ndb, int32_t min_total,
                       int32_t min_diag, int32_t w, int32_t max_cands, int32_t topk_a,
                       int32_t n_threads, int32_t* o_targets, int32_t* o_counts,
                       int32_t* o_scores, int32_t* o_diags) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < nq; i++) {
        int64_t base = (int64_t)i * max_cands;
        o_counts[i] = korc_query(qflat + qoff[i], qlen[i], k, alpha, koff, kent, kfreq, fthr, ndb,
                                 min_total, min_diag, w, topk_a, max_cands, o_targets + base,
                                 o_scores + base, o_diags + base);
    }
}

/*
 * BLOSUM-similar 5-mer scoring (kmer_index.py:1186-1330,
 * _score_query_similar_k5 / _batch_score_queries_similar): for every query
 * 5-mer with all residues < 20 and every 5-mer a with column score
 * S = sum_j sub[q_j][a_j] >= T and 0 < freq(a) <= freq_thresh, each
 * posting of a adds S to its target's count; targets with count >=
 * min_total_hits pass, in ascending id, or the top max_cands by (count
 * desc, id asc) when more pass (the reference's quicksort leaves the order
 * of equal counts unspecified).  Restated as a depth-first walk over the
 * five residues with suffix bounds (the best completion of a prefix is its
 * score plus the row maxima of the remaining query residues), which visits
 * exactly the codes with S >= T.
 */
typedef struct {
    const int8_t* sub;
    int32_t T;
    const int64_t* koff;
    const int64_t* kent;
    const int32_t* kfreq;
    int32_t fthr;
    int32_t* cnt;
    int32_t q[5];
    int32_t rest[6];   /* rest[d] = sum of row maxima of q[d..4] */
} ksim_ctx;

static void ksim_walk(ksim_ctx* c, int32_t depth, int32_t code, int32_t s) {
    if (depth == 5) {
        if (s < c->T) return;
        int32_t f = c->kfreq[code];
        if (f == 0 || f > c->fthr) return;
        for (int64_t h = c->koff[code]; h < c->koff[code + 1]; h++)
            c->cnt[(int32_t)(c->kent[h] >> 32)] += s;
        return;
    }
    for (int32_t a = 0; a < 20; a++) {
        int32_t s2 = s + c->sub[c->q[depth] * 20 + a];
        if (s2 + c->rest[depth + 1] < c->T) continue;
        ksim_walk(c, depth + 1, code * 20 + a, s2);
    }
}

static int32_t ksim_query(const uint8_t* q, int32_t qlen, const int64_t* koff,
                          const int64_t* kent, const int32_t* kfreq, int32_t fthr, int32_t ndb,
                          int32_t min_total, int32_t max_cands, const int8_t* sub, int32_t T,
                          int32_t* o_ids, int32_t* o_sc) {
    if (qlen < 5) return 0;
    int32_t rowmax[20];
    for (int32_t r = 0; r < 20; r++) {
        int32_t m = sub[r * 20];
        for (int32_t a = 1; a < 20; a++) if (sub[r * 20 + a] > m) m = sub[r * 20 + a];
        rowmax[r] = m;
    }
    ksim_ctx c;
    c.sub = sub; c.T = T; c.koff = koff; c.kent = kent; c.kfreq = kfreq; c.fthr = fthr;
    c.cnt = (int32_t*)calloc((size_t)(ndb > 0 ? ndb : 1), sizeof(int32_t));
    for (int32_t p = 0; p + 5 <= qlen; p++) {
        int ok = 1;
        for (int32_t j = 0; j < 5; j++) {
            c.q[j] = q[p + j];
            if (c.q[j] >= 20) ok = 0;
        }
        if (!ok) continue;
        c.rest[5] = 0;
        for (int32_t j = 4; j >= 0; j--) c.rest[j] = c.rest[j + 1] + rowmax[c.q[j]];
        ksim_walk(&c, 0, 0, 0);
    }
    korc_cand* cand = (korc_cand*)malloc(sizeof(korc_cand) * (size_t)(ndb > 0 ? ndb : 1));
    int32_t n = 0;
    for (int32_t t = 0; t < ndb; t++)
        if (c.cnt[t] >= min_total) {
            cand[n].id = t;
            cand[n].score = c.cnt[t];
            cand[n].diag = 0;
            n++;
        }
    n = korc_top(cand, n, max_cands);
    for (int32_t i = 0; i < n; i++) {
        o_ids[i] = cand[i].id;
        o_sc[i] = cand[i].score;
    }
    free(cand);
    free(c.cnt);
    return n;
}

void kmer_similar_oracle_batch(const uint8_t* qflat, const int64_t* qoff, const int32_t* qlen,
                               int32_t nq, const int64_t* koff, const int64_t* kent,
                               const int32_t* kfreq, int32_t fthr, int32_t ndb, int32_t min_total,
                               int32_t max_cands, const int8_t* sub, int32_t T, int32_t n_threads,
                               int32_t* o_targets, int32_t* o_counts, int32_t* o_scores) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < nq; i++) {
        int64_t base = (int64_t)i * max_cands;
        o_counts[i] = ksim_query(qflat + qoff[i], qlen[i], koff, kent, kfreq, fthr, ndb, min_total,
                                 max_cands, sub, T, o_targets + base, o_scores + base);
    }
}
