/*
 * CPU oracle for banded global (Needleman-Wunsch) identity -- TEST
 * INFRASTRUCTURE ONLY.  Nothing in the shipped package links or calls this
 * file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, as the checker.  It restates:
 *
 *   nw_oracle_pair    clustkit/pairwise.py:258-362 (_nw_identity) and
 *                     clustkit/pairwise.py:365-505 (_nw_identity_and_score)
 *   nw_oracle_batch   clustkit/pairwise.py:804-850 (_batch_align_compact) and
 *                     clustkit/pairwise.py:684-736 (_batch_align_compact_scored)
 *
 * Formulation (different from the reference's two toggled rows with border
 * patches): every row r has an explicit set of defined cells and any other
 * cell reads as H = -inf, E = -inf, Hm = Em = 0 ("-inf" is the reference's
 * -1,000,000 so that even values derived from it agree):
 *   row 0        columns 0..min(bw, lb): H = 0 at column 0, open + ext*(j-1)
 *   column 0     rows 1..bw+1:           H = open + ext*(r-1)
 *   r, j >= 1    |r - j| <= bw, j <= lb: the recurrence
 * Column 0 is defined one row further down than the band (row bw+1), which
 * is what the reference's left-border rule does (pairwise.py:311-314).  That
 * cell, like any cell just outside the band, can never change an output: the
 * diagonal candidate of its neighbour is always larger, so no test can tell
 * the two readings apart.
 * Recurrence and tie rules:
 *   E = extension if E_above + ext >= H_above + open, else opening
 *   F = extension if F_left + ext >= H_left + open, else opening; the first
 *       cell of a row has no extension candidate
 *   H = diagonal, replaced by E if strictly greater, then by F if strictly
 *       greater; each value carries the match count of the path it took
 * Early stop: after row r, if fl32(max row Hm + la - r) / fl32(la) < t the
 * result is identity 0 (score -1,000,000 in the scored form).
 * Scoring: plain form +2 / -1 on raw byte equality, gaps -10 / -1; scored
 * form BLOSUM-type table for codes < 20 and -4 otherwise, gaps -11 / -1;
 * matches count raw byte equality in both.
 *
 * Pinned by: the reference's own NW tests (tests/test_pairwise.py:71-142),
 * closed forms (identical sequences), an independent full-matrix Gotoh with
 * traceback under the same tie preferences on small cases, and brute force
 * over all global alignments of tiny sequences (tests/test_nw_oracle.py).
 */

#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define NW_NEG (-1000000)

static inline int32_t nw_sub(int scored, const int8_t* sub, uint8_t x, uint8_t y) {
    if (!scored) return x == y ? 2 : -1;
    return (x < 20 && y < 20) ? (int32_t)sub[x * 20 + y] : -4;
}

/* One pair, la <= lb.  ws: 8 rows of lb + 1 ints.  Returns identity; *score
 * receives H(la, lb) or -1,000,000 on an early stop. */
float nw_oracle_pair(const uint8_t* a, int32_t la, const uint8_t* b, int32_t lb,
                     int32_t band, float threshold, int32_t scored, const int8_t* sub,
                     int32_t* ws, int32_t* score) {
    if (la == 0 || lb == 0) {
        *score = 0;
        return 0.0f;
    }
    const int32_t longest = la > lb ? la : lb;
    const int32_t bw = (band <= 0 || longest <= 50) ? longest : band;
    const int32_t GO = scored ? -11 : -10, GE = -1;
    const int32_t n = lb + 1;
    int32_t *Hp = ws, *Mp = ws + n, *Ep = ws + 2 * n, *EMp = ws + 3 * n;
    int32_t *Hc = ws + 4 * n, *Mc = ws + 5 * n, *Ec = ws + 6 * n, *EMc = ws + 7 * n;

    /* Row 0. */
    int32_t plo = 0, phi = bw < lb ? bw : lb;
    for (int32_t j = 0; j <= phi; ++j) {
        Hp[j] = j == 0 ? 0 : GO + GE * (j - 1);
        Mp[j] = 0;
        Ep[j] = NW_NEG;
        EMp[j] = 0;
    }
#define PH(j) (((j) >= plo && (j) <= phi) ? Hp[j] : NW_NEG)
#define PM(j) (((j) >= plo && (j) <= phi) ? Mp[j] : 0)
#define PE(j) (((j) >= plo && (j) <= phi) ? Ep[j] : NW_NEG)
#define PEM(j) (((j) >= plo && (j) <= phi) ? EMp[j] : 0)

    for (int32_t r = 1; r <= la; ++r) {
        const int32_t lo = r - bw > 1 ? r - bw : 1;
        const int32_t hi = r + bw < lb ? r + bw : lb;
        const int32_t col0 = r <= bw + 1;               /* column 0 defined */
        const int32_t lo_def = col0 ? 0 : lo;
        if (col0) {
            Hc[0] = GO + GE * (r - 1);
            Mc[0] = 0;
            Ec[0] = NW_NEG;
            EMc[0] = 0;
        }
        int32_t F = NW_NEG, FM = 0, max_m = 0;
        for (int32_t j = lo; j <= hi; ++j) {
            const uint8_t x = a[r - 1], y = b[j - 1];
            const int32_t s = nw_sub(scored, sub, x, y), mt = x == y;
            const int32_t d = PH(j - 1) + s, dm = PM(j - 1) + mt;
            const int32_t eext = PE(j) + GE, eop = PH(j) + GO;
            int32_t e, em;
            if (eext >= eop) { e = eext; em = PEM(j); } else { e = eop; em = PM(j); }
            const int32_t left_h = (j - 1 >= lo_def) ? Hc[j - 1] : NW_NEG;
            const int32_t left_m = (j - 1 >= lo_def) ? Mc[j - 1] : 0;
            const int32_t fext = j > lo ? F + GE : NW_NEG, fop = left_h + GO;
            if (fext >= fop) { F = fext; } else { F = fop; FM = left_m; }
            int32_t h = d, hm = dm;
            if (e > h) { h = e; hm = em; }
            if (F > h) { h = F; hm = FM; }
            Hc[j] = h;
            Mc[j] = hm;
            Ec[j] = e;
            EMc[j] = em;
            if (hm > max_m) max_m = hm;
        }
        int32_t* t;
        t = Hp; Hp = Hc; Hc = t;
        t = Mp; Mp = Mc; Mc = t;
        t = Ep; Ep = Ec; Ec = t;
        t = EMp; EMp = EMc; EMc = t;
        plo = lo_def;
        phi = hi;
        if ((float)(max_m + (la - r)) / (float)la < threshold) {
            *score = NW_NEG;
            return 0.0f;
        }
    }
    *score = PH(lb);
    return (float)PM(lb) / (float)la;
#undef PH
#undef PM
#undef PE
#undef PEM
}

/* Batch form: length-ratio test, length-difference test, adaptive band
 * min(bw, max(10, diff + 10)), shorter sequence as rows, mask = identity >= t.
 * out_scores may be NULL (plain form). */
void nw_oracle_batch(const int32_t* pairs, int64_t m, const uint8_t* flat, const int64_t* offsets,
                     const int32_t* lengths, float threshold, int32_t band_width,
                     int32_t scored, const int8_t* sub, int32_t n_threads, float* out_sims,
                     int32_t* out_scores, uint8_t* out_mask) {
    int32_t maxlen = 0;
    for (int64_t p = 0; p < m; ++p) {
        const int32_t x = lengths[pairs[2 * p]], y = lengths[pairs[2 * p + 1]];
        if (x > maxlen) maxlen = x;
        if (y > maxlen) maxlen = y;
    }
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
    {
        int32_t* ws = (int32_t*)malloc(sizeof(int32_t) * 8 * ((size_t)maxlen + 1));
#pragma omp for schedule(dynamic, 64)
        for (int64_t p = 0; p < m; ++p) {
            const int32_t i = pairs[2 * p], j = pairs[2 * p + 1];
            const int32_t li = lengths[i], lj = lengths[j];
            const int32_t sh = li < lj ? li : lj, lg = li > lj ? li : lj;
            float sim = 0.0f;
            int32_t sc = NW_NEG;
            int32_t pass = 0;
            const int32_t diff = li > lj ? li - lj : lj - li;
            if (lg > 0 && (float)sh / (float)lg < threshold) {
                pass = 0;
            } else if (diff > band_width) {
                pass = 0;
            } else {
                int32_t ab = diff + 10 > 10 ? diff + 10 : 10;
                if (band_width < ab) ab = band_width;
                if (li <= lj)
                    sim = nw_oracle_pair(flat + offsets[i], li, flat + offsets[j], lj, ab,
                                         threshold, scored, sub, ws, &sc);
                else
                    sim = nw_oracle_pair(flat + offsets[j], lj, flat + offsets[i], li, ab,
                                         threshold, scored, sub, ws, &sc);
                pass = sim >= threshold;
            }
            out_sims[p] = sim;
            if (out_scores) out_scores[p] = sc;
            out_mask[p] = (uint8_t)pass;
        }
        free(ws);
    }
}
