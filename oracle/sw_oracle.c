/*
 * CPU oracle for the alignment stage -- TEST INFRASTRUCTURE ONLY.
 *
 * Nothing in the shipped package links or calls this file. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only
 * as the checker. It restates, in its own formulation, the semantics of:
 *
 *   sw_oracle_pair      clustkit/csrc/sw_align.c:25-177  (sw_align_one)
 *   sw_oracle_batch     clustkit/csrc/sw_align.c:182-266 (batch_sw_align_c)
 *   jaccard_oracle_*    clustkit/pairwise.py:1564-1612   (_jaccard_sorted,
 *                                                         _batch_jaccard)
 *
 * Formulation used here (deliberately different from the reference's
 * row-toggling buffers with border patches): each DP row i has an explicit
 * column interval [lo_i, hi_i]; any cell outside the interval of its row is
 * "outside the band" and reads as H = 0, Hm = 0, E = -inf, Em = 0.  Column 0
 * is outside every band.  The tie rules are those of the reference:
 *   E, F  : extension wins ties against opening            (sw_align.c:129,141)
 *   H     : first strictly greater of 0, diag, E, F         (sw_align.c:147-151)
 *   max   : first cell in row-major order with the maximum (sw_align.c:156-159)
 * Parity of this restatement is pinned by golden vectors produced from the
 * reference kernel compiled from its own sources (tests/golden), by brute
 * force over all alignments of tiny sequences and by closed forms.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -shared -fPIC).
 */

#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_NEG (-1000000)

typedef struct {
    int32_t score;    /* best local score, 0 when no positive cell        */
    int32_t matches;  /* identical residues along the path to that cell   */
    int32_t end_i;    /* 1-based row of the first best cell (0 if none)   */
    int32_t end_j;    /* 1-based column of the first best cell            */
    float identity;   /* matches / min(len_a, len_b), 0 when score <= 0   */
} orc_result;

static inline int32_t orc_sub(const int8_t* sub, uint8_t x, uint8_t y) {
    return (x < 20 && y < 20) ? (int32_t)sub[x * 20 + y] : -4;
}

/* One pair, len_a <= len_b expected by the caller (the batch swaps). */
static orc_result orc_align(const uint8_t* a, int32_t la,
                            const uint8_t* b, int32_t lb,
                            int32_t band, int32_t hint,
                            const int8_t* sub, int32_t* ws) {
    orc_result r = {0, 0, 0, 0, 0.0f};
    if (la <= 0 || lb <= 0) return r;
    int32_t longest = la > lb ? la : lb;
    int32_t bw = (band <= 0 || longest <= 50) ? longest : band;

    /* ws holds 8 rows of lb+1 ints: prev/cur for H, Hm, E, Em. */
    int32_t n = lb + 1;
    int32_t *Hp = ws, *Mp = ws + n, *Ep = ws + 2 * n, *EMp = ws + 3 * n;
    int32_t *Hc = ws + 4 * n, *Mc = ws + 5 * n, *Ec = ws + 6 * n, *EMc = ws + 7 * n;
    int32_t plo = 1, phi = 0; /* previous row interval: empty */

    for (int32_t i = 1; i <= la; i++) {
        int32_t lo = i + hint - bw, hi = i + hint + bw;
        if (lo < 1) lo = 1;
        if (hi > lb) hi = lb;
        if (lo > hi) { plo = 1; phi = 0; continue; }
        uint8_t ra = a[i - 1];
        int32_t f = ORC_NEG, fm = 0;          /* F of the cell to the left   */
        int32_t hl = 0, hlm = 0;              /* H of the cell to the left   */
        for (int32_t j = lo; j <= hi; j++) {
            int in_up = (j >= plo && j <= phi);
            int in_dg = (j - 1 >= plo && j - 1 <= phi);
            int32_t hu = in_up ? Hp[j] : 0, hum = in_up ? Mp[j] : 0;
            int32_t eu = in_up ? Ep[j] : ORC_NEG, eum = in_up ? EMp[j] : 0;
            int32_t hd = in_dg ? Hp[j - 1] : 0, hdm = in_dg ? Mp[j - 1] : 0;
            uint8_t rb = b[j - 1];

            int32_t dv = hd + orc_sub(sub, ra, rb);
            int32_t dm = hdm + (ra == rb);

            int32_t ev, em;
            if (eu - 1 >= hu - 11) { ev = eu - 1; em = eum; }
            else                   { ev = hu - 11; em = hum; }

            int32_t fext = (j == lo) ? ORC_NEG : f - 1;
            if (fext >= hl - 11) { f = fext; /* fm unchanged */ }
            else                 { f = hl - 11; fm = hlm; }

            int32_t h = 0, hm = 0;
            if (dv > h) { h = dv; hm = dm; }
            if (ev > h) { h = ev; hm = em; }
            if (f > h)  { h = f;  hm = fm; }

            Hc[j] = h; Mc[j] = hm; Ec[j] = ev; EMc[j] = em;
            hl = h; hlm = hm;
            if (h > r.score) { r.score = h; r.matches = hm; r.end_i = i; r.end_j = j; }
        }
        int32_t* t;
        t = Hp; Hp = Hc; Hc = t;
        t = Mp; Mp = Mc; Mc = t;
        t = Ep; Ep = Ec; Ec = t;
        t = EMp; EMp = EMc; EMc = t;
        plo = lo; phi = hi;
    }
    if (r.score <= 0) {
        r.score = 0; r.matches = 0; r.end_i = 0; r.end_j = 0; r.identity = 0.0f;
    } else {
        r.identity = (float)r.matches / (float)la;
    }
    return r;
}

/* Single pair with the batch's orientation rules applied (swap so the
 * shorter sequence is the row sequence, clamp and negate the hint). */
void sw_oracle_pair(const uint8_t* si, int32_t li, const uint8_t* sj, int32_t lj,
                    int32_t band, int32_t hint, const int8_t* sub,
                    int32_t* out5 /* score, matches, end_i, end_j, swapped */,
                    float* out_identity) {
    orc_result r = {0, 0, 0, 0, 0.0f};
    int swapped = 0;
    if (li > 0 && lj > 0) {
        int32_t sh = li < lj ? li : lj;
        if (hint > sh || hint < -sh) hint = 0;
        int32_t lb = li > lj ? li : lj;
        int32_t* ws = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)(lb + 1));
        if (li <= lj) r = orc_align(si, li, sj, lj, band, hint, sub, ws);
        else { r = orc_align(sj, lj, si, li, band, -hint, sub, ws); swapped = 1; }
        free(ws);
    }
    out5[0] = r.score; out5[1] = r.matches; out5[2] = r.end_i; out5[3] = r.end_j;
    out5[4] = swapped;
    *out_identity = r.identity;
}

/* Batch over candidate pairs; out_mask = score > 0 as in sw_align.c:260.
 * out_matches / out_end may be NULL. */
void sw_oracle_batch(const int32_t* pairs, const uint8_t* flat,
                     const int64_t* offsets, const int32_t* lengths,
                     int64_t M, int32_t band, const int8_t* sub,
                     const int32_t* hints, int32_t n_threads,
                     float* out_sims, int32_t* out_scores, uint8_t* out_mask,
                     int32_t* out_matches, int32_t* out_end /* 2*M or NULL */) {
    int32_t lmax = 0;
    for (int64_t p = 0; p < M; p++) {
        int32_t x = lengths[pairs[2 * p]], y = lengths[pairs[2 * p + 1]];
        if (x > lmax) lmax = x;
        if (y > lmax) lmax = y;
    }
    if (n_threads > 0) omp_set_num_threads(n_threads);
    #pragma omp parallel
    {
        int32_t* ws = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)(lmax + 2));
        #pragma omp for schedule(dynamic, 16)
        for (int64_t p = 0; p < M; p++) {
            int32_t x = pairs[2 * p], y = pairs[2 * p + 1];
            int32_t lx = lengths[x], ly = lengths[y];
            orc_result r = {0, 0, 0, 0, 0.0f};
            if (lx > 0 && ly > 0) {
                int32_t h = hints ? hints[p] : 0;
                int32_t sh = lx < ly ? lx : ly;
                if (h > sh || h < -sh) h = 0;
                if (lx <= ly) r = orc_align(flat + offsets[x], lx, flat + offsets[y], ly, band, h, sub, ws);
                else          r = orc_align(flat + offsets[y], ly, flat + offsets[x], lx, band, -h, sub, ws);
            }
            out_sims[p] = r.identity;
            out_scores[p] = r.score;
            out_mask[p] = r.score > 0;
            if (out_matches) out_matches[p] = r.matches;
            if (out_end) { out_end[2 * p] = r.end_i; out_end[2 * p + 1] = r.end_j; }
        }
        free(ws);
    }
}

/* Merge-join estimate over two sorted bottom-s sketches
 * (pairwise.py:1564-1595): walk both arrays; stop after s union steps, at
 * the end of either array, or when both heads are the all-ones sentinel. */
float jaccard_oracle_one(const uint64_t* x, const uint64_t* y, int32_t s) {
    const uint64_t SENT = ~(uint64_t)0;
    int32_t ix = 0, iy = 0, uni = 0, both = 0;
    while (uni < s && ix < s && iy < s) {
        uint64_t u = x[ix], v = y[iy];
        if (u == SENT && v == SENT) break;
        if (u == v) { both++; ix++; iy++; }
        else if (u < v) ix++;
        else iy++;
        uni++;
    }
    if (uni == 0) return 0.0f;
    return (float)((double)both / (double)uni);
}

void jaccard_oracle_batch(const int32_t* pairs, const uint64_t* sketches,
                          int64_t M, int32_t s, int32_t n_threads, float* out) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < M; p++)
        out[p] = jaccard_oracle_one(sketches + (int64_t)pairs[2 * p] * s,
                                    sketches + (int64_t)pairs[2 * p + 1] * s, s);
}
