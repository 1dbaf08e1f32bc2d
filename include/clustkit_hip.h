/*
 * clustkit_hip.h -- C-ABI of the MI355X alignment stage (libclustkit_hip.so).
 *
 * Plain pointers and sizes only.  Every entry point returns CK_OK (0) or a
 * negative CK_ERR_* code; ck_last_error() gives a message for the calling
 * thread.  "d_" pointers are device (HBM) pointers, "h_" pointers are host
 * pointers.  Calls taking a `stream` (a hipStream_t, NULL = default stream)
 * are asynchronous with respect to the host unless stated otherwise.
 *
 * Scoring is fixed to the reference's: BLOSUM62-style 20x20 int8 matrix in
 * alphabet order ACDEFGHIKLMNPQRSTVWY, codes >= 20 score -4, gap open -11,
 * gap extend -1 (clustkit/csrc/sw_align.c:45-47,116-119).
 */
#ifndef CLUSTKIT_HIP_H
#define CLUSTKIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CK_OK               0
#define CK_ERR_INVALID     -1  /* bad argument (null pointer, negative size...) */
#define CK_ERR_HIP         -2  /* HIP runtime error                              */
#define CK_ERR_NOMEM       -3  /* device allocation failed                        */
#define CK_ERR_NODEVICE    -4  /* no GPU visible                                  */
#define CK_ERR_UNSUPPORTED -5  /* combination not supported by this build         */

/* Flags for ck_sw_align_device / ck_sw_align_host. */
#define CK_SW_LENGTH_RATIO  0x1u  /* skip pairs with fl32(short)/fl32(long) <
                                     fl32(threshold): identity 0, score 0, mask 0
                                     (clustkit/pairwise.py:1243-1252)          */
#define CK_SW_MASK_IDENTITY 0x2u  /* out_mask = identity >= threshold instead of
                                     score > 0 (pairwise.py:107 / 1394-1396)   */
#define CK_SW_PLAIN         0x4u  /* run every pair through the plain kernel
                                     (one pair per thread, int32, one launch per
                                     pair batch) instead of the batched kernel  */
#define CK_SW_NO_GRAPH      0x8u  /* launch kernels directly, no hipGraph replay */
#define CK_SW_TIMING        0x10u /* record HIP events around the call's phases on
                                     `stream` (read with ck_sw_last_timing)     */
#define CK_SW_LATENCY       0x20u /* one pair per wavefront (its 64 lanes one
                                     chain over stripes of 64 rows) instead of
                                     four: ~4x shorter time per pair, for small
                                     batches that cannot fill the GPU; same
                                     results                                    */

/* Flags for ck_nw_align_device / ck_nw_align_host. */
#define CK_NW_PLAIN         0x4u  /* every pair through the one-pair-per-thread
                                     kernel instead of the 16-lane kernel     */
#define CK_NW_RAW           0x8u  /* the inner function alone (_nw_identity,
                                     pairwise.py:258-287): no length tests, the
                                     band as given, shorter sequence on rows  */
#define CK_NW_STRIPE        0x10u /* packed-key pairs one stripe at a time
                                     instead of stripe-pipelined; same results */

/* Last error message of the calling thread (never NULL). */
const char* ck_last_error(void);

/* Library version string. */
const char* ck_version(void);

/* Number of visible GPUs (0 when none); never fails. */
int ck_device_count(void);

/* Select the current device for the calling thread and upload the constant
 * tables.  Called implicitly by the other entry points (device 0). */
int ck_init(int device);

/* Properties of `device` from hipGetDeviceProperties: compute units, peak
 * engine clock (kHz), HBM capacity (bytes) and the architecture name
 * (e.g. "gfx950:sramecc+:xnack-", truncated to arch_name_len - 1 characters).
 * NULL pointers are skipped.  No reference counterpart: the benchmark
 * derives the VALU peak (CUs x 64 lanes x clock) from it. */
int ck_device_props(int device, int32_t* num_cu, int32_t* clock_khz, int64_t* hbm_bytes,
                    char* arch_name, int32_t arch_name_len);

/* Batched-kernel configuration, process-wide; results do not depend on it
 * (no reference counterpart: the knobs of the autotuner,
 * clustkit_amd/autotune.py).  lanes_per_pair: 16 (four pairs per wavefront),
 * 8 (eight pairs per wavefront) or 0 (8 for bands 1..32, else 16);
 * bin_shift: the work queue orders pairs by estimated steps >> bin_shift
 * (0..12, default 5); pipelined: 1 (default) runs 16- and 8-lane segments
 * with the stripe-pipelined kernel (each lane walks its stripes back to
 * back), 0 with one stripe at a time. */
int ck_sw_set_config(int32_t lanes_per_pair, int32_t bin_shift, int32_t pipelined);
int ck_sw_get_config(int32_t* lanes_per_pair, int32_t* bin_shift, int32_t* pipelined);

/* Drop-in for the reference C kernel
 *   batch_sw_align_c(pairs, flat_sequences, offsets, lengths, M, band_width,
 *                    sub_matrix, threshold, diag_hints, out_sims,
 *                    out_scores, out_mask)
 * (clustkit/csrc/sw_align.c:182-195, ctypes binding pairwise.py:50-58 and
 * kmer_index.py:2235-2243).  Host pointers, synchronous, same outputs:
 * out_sims = matches / min(len), out_scores = best local score (0 if none),
 * out_mask = score > 0.  `threshold` is unused, as in the reference.
 * The number of sequences is taken as 1 + the largest index in `pairs`. */
int ck_batch_sw_align(const int32_t* pairs, const uint8_t* flat_sequences,
                      const int64_t* offsets, const int32_t* lengths, int32_t M,
                      int32_t band_width, const int8_t* sub_matrix, float threshold,
                      const int32_t* diag_hints, float* out_sims, int32_t* out_scores,
                      uint8_t* out_mask);

/* Host-pointer form with explicit sizes and flags (used by the Python mirror
 * of clustkit/pairwise.py:_batch_sw_c / _batch_sw_compact_scored).  Copies
 * inputs to HBM, runs ck_sw_align_device, copies results back; synchronous.
 * out_matches may be NULL.  Host-pointer calls (this one, ck_nw_align_host,
 * ck_jaccard_host, ck_graph_csr_host, ck_batch_score_queries,
 * ck_batch_sw_align) share per-device staging buffers: calls from several
 * host threads on one device are serialised inside the library, each from
 * its copy-in to its copy-out. */
int ck_sw_align_host(const int32_t* h_pairs, int64_t M, const uint8_t* h_flat,
                     int64_t flat_len, const int64_t* h_offsets,
                     const int32_t* h_lengths, int64_t n_seqs, int32_t band_width,
                     const int8_t* h_sub_matrix, const int32_t* h_diag_hints,
                     float threshold, uint32_t flags, float* h_out_sims,
                     int32_t* h_out_scores, int32_t* h_out_matches,
                     uint8_t* h_out_mask);

/* Device-resident form: every array already in HBM.  The batched kernel is
 * a persistent launch over a device work queue ordered by length-product
 * bins; pairs whose int16 score lane saturates (or whose shorter sequence
 * exceeds 8191 residues, or whose band exceeds the LDS row buffer) are
 * re-run by the int32 kernel inside the same call.  `max_len` is an upper
 * bound of lengths[] (sizes the int32 kernel's workspace).  Workspace is
 * owned by the library (grown on demand, reused across calls).  The call
 * returns once all work is enqueued on `stream`. */
int ck_sw_align_device(const int32_t* d_pairs, int64_t M, const uint8_t* d_flat,
                       const int64_t* d_offsets, const int32_t* d_lengths,
                       int32_t max_len, int32_t band_width, const int8_t* h_sub_matrix,
                       const int32_t* d_diag_hints, float threshold, uint32_t flags,
                       float* d_out_sims, int32_t* d_out_scores,
                       int32_t* d_out_matches, uint8_t* d_out_mask, void* stream);

/* Same call, same kernels and same results as ck_sw_align_device, without
 * any device allocation: no counterpart in the reference.  For callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its work memory is one fixed 48 MiB
 * array per GPU inside the library (allocated when the library's code is
 * loaded on that GPU, held until the process ends); the caller passes
 * inputs and outputs as for ck_sw_align_device.  Limits, rejected with
 * CK_ERR_INVALID before any device work: M <= 1,048,576 pairs; one-pair-per-
 * thread band rows of at most 8,192 entries (band_width <= 4,095, or with
 * band_width <= 0, max_len <= 4,095); CK_SW_TIMING not accepted.  Launches
 * directly (no hipGraph).  Every call on one GPU uses that GPU's array: calls
 * that overlap on one GPU (several streams or threads) race; serialise them.
 * Call ck_init(device) once beforehand so that the first call does not
 * create the library's per-device context. */
int ck_sw_align_device_noalloc(const int32_t* d_pairs, int64_t M, const uint8_t* d_flat,
                               const int64_t* d_offsets, const int32_t* d_lengths,
                               int32_t max_len, int32_t band_width,
                               const int8_t* h_sub_matrix, const int32_t* d_diag_hints,
                               float threshold, uint32_t flags, float* d_out_sims,
                               int32_t* d_out_scores, int32_t* d_out_matches,
                               uint8_t* d_out_mask, void* stream);

/* Statistics of the last ck_sw_align_device call on this thread (after the
 * stream has been synchronised): pairs queued for the 16-bit batched
 * kernel, pairs it could not finish or could not take (re-run: wide-score
 * pass or int32 kernel), pairs skipped by the length-ratio filter. */
int ck_sw_last_stats(int64_t* batched_pairs, int64_t* rerun_pairs, int64_t* skipped_pairs);

/* How the last call's pairs that left the 16-bit batched kernel were
 * finished: wide_pairs went through the wide-score (17-bit) wavefront pass,
 * int32_pairs through the one-pair-per-thread int32 kernel (rows > 8191,
 * bands wider than 511 positions, scores >= 65500, raw codes > 30).  No
 * reference counterpart (diagnostics of this implementation). */
int ck_sw_last_paths(int64_t* wide_pairs, int64_t* int32_pairs);

/* Status of the last ck_sw_align_device / ck_nw_align_device call on this
 * thread (waits for its counters): CK_ERR_INVALID when a pair's band row was
 * wider than the one-pair-per-thread workspace, i.e. `max_len` was below a
 * pair's longer sequence with band_width <= 0 (that pair's outputs are 0);
 * CK_OK otherwise.  The host-pointer forms report the same condition as
 * their return value.  No reference counterpart (the reference does not
 * check its workspace size). */
int ck_sw_last_status(void);

/* Event timings (milliseconds) of the last ck_sw_align_device call made with
 * CK_SW_TIMING on this thread: queue construction (reset, prep, scan,
 * scatter), main alignment kernel(s), and the int32 re-run.  Waits for the
 * last event.  Pass NULL for values not wanted. */
int ck_sw_last_timing(float* queue_ms, float* main_ms, float* rerun_ms);

/* ---- banded global alignment (NW identity) ------------------------------
 *
 * Replaces the Numba batch functions of clustkit/pairwise.py:
 *   _batch_align_compact (804-850) over _nw_identity (258-362): pass
 *     h_sub_matrix = NULL -> +2 / -1 on byte equality, gaps -10 / -1;
 *     out_scores may be NULL;
 *   _batch_align_compact_scored (684-736) over _nw_identity_and_score
 *     (365-505): pass the 20x20 table -> table score for codes < 20, -4
 *     otherwise, gaps -11 / -1; out_scores = H(la, lb), or -1,000,000 for
 *     pairs removed by the length tests or stopped early.
 * Per pair: fl32(short)/fl32(long) < threshold or |len_i - len_j| >
 * band_width -> identity 0, mask 0; else the adaptive half band
 * min(band_width, max(10, diff + 10)), the full matrix when that is <= 0 or
 * the longer sequence has <= 50 residues; early stop (identity 0) after any
 * row whose best match count plus the remaining rows is below threshold *
 * shorter; identity = fl32(matches at (la, lb)) / fl32(shorter); mask =
 * identity >= threshold.  Called by clustkit/search.py:668-685 (LSH search),
 * clustkit/kmer_index.py:2281-2300 (k-mer search with local_alignment=False),
 * clustkit/pairwise.py:1031-1038 (refinement of borderline SW hits) and
 * pairwise.py:1197-1206 (use_sw=False).  The reference has no C ABI for these
 * (Numba only); the signature follows ck_sw_align_device.
 * After a call, the second value of ck_sw_last_paths is the number of pairs
 * run one per thread (bands wider than 479 positions).  `max_len` must bound lengths[] (it sizes
 * the one-pair-per-thread workspace; the host form reports a violation). */
int ck_nw_align_device(const int32_t* d_pairs, int64_t M, const uint8_t* d_flat,
                       const int64_t* d_offsets, const int32_t* d_lengths,
                       int32_t max_len, int32_t band_width, const int8_t* h_sub_matrix,
                       float threshold, uint32_t flags, float* d_out_sims,
                       int32_t* d_out_scores, uint8_t* d_out_mask, void* stream);

/* Host-pointer form (synchronous), used by the Python mirror of the batch
 * functions above. */
int ck_nw_align_host(const int32_t* h_pairs, int64_t M, const uint8_t* h_flat,
                     int64_t flat_len, const int64_t* h_offsets,
                     const int32_t* h_lengths, int64_t n_seqs, int32_t band_width,
                     const int8_t* h_sub_matrix, float threshold, uint32_t flags,
                     float* h_out_sims, int32_t* h_out_scores, uint8_t* h_out_mask);

/* ---- similarity graph (CSR) ------------------------------------------------
 *
 * Replaces clustkit/graph.py:7-39 (build_similarity_graph): the (N, N)
 * symmetric float32 CSR of the passing pairs -- entries (i, j, w) and
 * (j, i, w) per pair, column indices sorted within each row, duplicate
 * entries summed, explicit zeros kept, as scipy's coo -> csr conversion
 * gives them.  d_indptr holds N + 1 entries, d_indices / d_data room for
 * 2K; *d_nnz receives the entry count.  Pair indices must lie in [0, N)
 * (the host form checks them).  Called by clustkit/pipeline.py:118 and
 * pipeline.py:696 on the passing pairs of the alignment stage. */
int ck_graph_csr_device(const int32_t* d_pairs, int64_t K, const float* d_sims,
                        int64_t N, int32_t* d_indptr, int32_t* d_indices, float* d_data,
                        int64_t* d_nnz, void* stream);

/* Host-pointer form (synchronous); h_indices / h_data need room for 2K. */
int ck_graph_csr_host(const int32_t* h_pairs, int64_t K, const float* h_sims, int64_t N,
                      int32_t* h_indptr, int32_t* h_indices, float* h_data,
                      int64_t* h_nnz);

/* ---- k-mer inverted-index scoring (search stage 1) ----------------------
 * Drop-in for the reference C kernel
 *   batch_score_queries_c(q_flat, q_offsets, q_lengths, nq, k, alpha_size,
 *       kmer_offsets, kmer_entries, compact_entries, kmer_freqs, freq_thresh,
 *       num_db, min_total_hits, min_diag_hits, diag_bin_width, max_cands,
 *       phase_a_topk, out_targets, out_counts, out_scores, out_diags)
 * (clustkit/csrc/kmer_score.c:420-444; ctypes binding kmer_index.py:1582-1595).
 * Host pointers, synchronous.  Phase A hit counts saturate at 32767; top-K
 * ties break by ascending target id (kmer_score.c:143-171); rows hold
 * out_counts[q] entries, in ascending id when not truncated, else by
 * (score desc, id asc).  out_diags is 0 when min_diag_hits <= 1 (the
 * reference leaves it unwritten then).  compact_entries may be NULL. */
int ck_batch_score_queries(const uint8_t* q_flat, const int64_t* q_offsets,
                           const int32_t* q_lengths, int32_t nq, int32_t k, int32_t alpha_size,
                           const int64_t* kmer_offsets, const int64_t* kmer_entries,
                           const int32_t* compact_entries, const int32_t* kmer_freqs,
                           int32_t freq_thresh, int32_t num_db, int32_t min_total_hits,
                           int32_t min_diag_hits, int32_t diag_bin_width, int32_t max_cands,
                           int32_t phase_a_topk, int32_t* out_targets, int32_t* out_counts,
                           int32_t* out_scores, int32_t* out_diags);

/* Device-resident form over an index held in HBM: postings split into
 * int32 target ids and int32 positions; d_tile_off (nullable) is the
 * per-(k-mer, 8192-target tile) offset table of ck_kmer_tile_offsets_host,
 * which switches Phase A/B to LDS-tiled counting (dense indexes); without
 * it counting uses global atomics.  d_db_lengths bounds the diagonal bins.
 * d_db_flat / d_db_offsets (both or neither; db_flat_len bytes) are the
 * database residues the index was built from: with them, code spaces of at
 * most 8192 codes (alpha^k, k <= 5, alpha <= 128) and max_qlen <= 16384, Phase B walks the
 * survivors' own windows against an LDS table of the query's k-mer
 * positions instead of re-reading the query's posting lists (same hits,
 * same results).  d_db_codes / d_db_code_off (both or neither; the output of
 * ck_kmer_db_codes_device, 16-byte aligned) give that Phase B the windows'
 * codes directly (any alphabet up to 8192 codes, with or without the
 * residues).  max_qlen must bound every query length of the batch.
 * Queries are processed in chunks that fit the workspace budget.  Outputs
 * are [nq * max_cands] rows and [nq] counts.  Asynchronous on `stream`. */
int ck_kmer_score_device(const uint8_t* d_qflat, int64_t qflat_len, const int64_t* d_qoff,
                         const int32_t* d_qlen, int32_t nq, int32_t max_qlen, int32_t k,
                         int32_t alpha, const int64_t* d_kmer_offsets,
                         const int32_t* d_post_tid, const int32_t* d_post_pos,
                         const int32_t* d_kmer_freqs, const int64_t* d_tile_off,
                         const int32_t* d_db_lengths, const uint8_t* d_db_flat,
                         int64_t db_flat_len, const int64_t* d_db_offsets,
                         const int16_t* d_db_codes, const int64_t* d_db_code_off, int32_t num_db,
                         int32_t max_db_len, int32_t freq_thresh, int32_t min_total_hits,
                         int32_t min_diag_hits, int32_t diag_bin_width, int32_t max_cands,
                         int32_t phase_a_topk, int32_t* d_out_targets, int32_t* d_out_counts,
                         int32_t* d_out_scores, int32_t* d_out_diags, void* stream);

/* BLOSUM-similar k-mer scoring (kmer_index.py:1186-1330: _score_query_similar_k5
 * and _batch_score_queries_similar, the reference's stage 1 for k >= 5 protein
 * indexes with kmer_score_thresh > 0) over a k = 5, alphabet-20 index in HBM.
 * For every query 5-mer with all residues < 20 and every 5-mer a whose
 * column score S = sum_j sub[q_j][a_j] >= kmer_score_thresh (>= 1) with
 * 0 < freq(a) <= freq_thresh, each posting of a adds S to its target's
 * count; the targets with count >= min_total_hits are the candidates, in
 * ascending id, or, when there are more than max_cands, the max_cands with
 * the highest counts in descending count (ties in ascending id; the
 * reference's quicksort leaves tie order unspecified).  h_sub_matrix:
 * host int8[400], row = query residue.  Outputs as ck_kmer_score_device:
 * out_scores holds the counts (saturated at 32767, the selection's scale),
 * out_diags 0.  d_db_lengths bounds nothing here but is validated as for
 * ck_kmer_score_device.  Asynchronous on `stream`. */
int ck_kmer_score_similar_device(const uint8_t* d_qflat, int64_t qflat_len,
                                 const int64_t* d_qoff, const int32_t* d_qlen, int32_t nq,
                                 int32_t max_qlen, const int64_t* d_kmer_offsets,
                                 const int32_t* d_post_tid, const int32_t* d_kmer_freqs,
                                 const int32_t* d_db_lengths, int32_t num_db,
                                 int32_t max_db_len, const int8_t* h_sub_matrix,
                                 int32_t kmer_score_thresh, int32_t freq_thresh,
                                 int32_t min_total_hits, int32_t max_cands,
                                 int32_t* d_out_targets, int32_t* d_out_counts,
                                 int32_t* d_out_scores, int32_t* d_out_diags, void* stream);

/* Search stage-2 input (no reference counterpart as a C function; it is
 * the flattening loop of kmer_index.py:1718-1736 on device):
 * d_pairs[2*(q*max_cands + c)] = (q_index_base + q, d_targets[q*max_cands + c])
 * for c < d_counts[q], else (empty_index, empty_index) -- an index of a
 * zero-length sequence, which ck_sw_align_device answers with zeros. */
int ck_search_pairs_device(const int32_t* d_targets, const int32_t* d_counts, int32_t nq,
                           int32_t max_cands, int32_t q_index_base, int32_t empty_index,
                           int32_t* d_pairs, void* stream);

/* Tile offset table (host arrays): out[c*(T+1) + j] = first posting of code
 * c whose target id is >= j*8192, T = ceil(num_db / 8192). */
int ck_kmer_tile_offsets_host(const int64_t* kmer_offsets, const int32_t* post_tid,
                              int64_t n_codes, int32_t num_db, int64_t* out);

/* Device build of the CSR k-mer index of a sequence table in HBM; replaces
 * build_kmer_index (clustkit/kmer_index.py:117-162, called by
 * clustkit/database.py:127-136 when a database is built).  Windows whose k
 * residues are all < alpha are indexed under code sum(r_j * alpha^(k-1-j));
 * postings of a code are in ascending (sequence, position) order, split
 * into int32 target ids and positions.  Outputs: d_kmer_offsets int64
 * [alpha^k + 1], d_kmer_freqs int32[alpha^k], d_post_tid / d_post_pos int32
 * [cap]; cap >= sum over sequences of max(len - k + 1, 0) always suffices
 * (CK_ERR_INVALID, with *out_total set, when it is too small).  Allocates
 * and frees its own sort buffers (about 24 bytes per window) and returns
 * after `stream` has drained; *out_total = postings written. */
int ck_kmer_index_build_device(const uint8_t* d_flat, const int64_t* d_offsets,
                               const int32_t* d_lengths, int32_t n_seqs, int32_t k, int32_t alpha,
                               int64_t cap, int64_t* d_kmer_offsets, int32_t* d_kmer_freqs,
                               int32_t* d_post_tid, int32_t* d_post_pos, int64_t* out_total,
                               void* stream);

/* Window codes of a database for the target-major Phase B (no reference
 * counterpart: the codes are those the index build posts,
 * kmer_index.py:117-162, kept per window instead).  Window p of sequence s
 * (code = its k residues in base alpha, first most significant) stores
 * code + 1 at d_codes[d_code_off[s] + p]; windows holding a residue >= alpha,
 * and the padding after the last window up to d_code_off[s + 1], store 0.  The
 * caller sizes the run of sequence s as max(0, len - k + 1) rounded up to a
 * multiple of 8 (d_code_off int64[n_seqs + 1], exclusive prefix sums).
 * alpha^k <= 8192, k <= 5; d_codes 16-byte aligned.  Asynchronous. */
int ck_kmer_db_codes_device(const uint8_t* d_flat, const int64_t* d_offsets,
                            const int32_t* d_lengths, int32_t n_seqs, int32_t k, int32_t alpha,
                            const int64_t* d_code_off, int16_t* d_codes, void* stream);

/* Device form of ck_kmer_tile_offsets_host: d_out int64[n_codes * (nt + 1)],
 * nt = ceil(num_db / 8192).  Asynchronous on `stream`. */
int ck_kmer_tile_offsets_device(const int64_t* d_kmer_offsets, const int32_t* d_post_tid,
                                int64_t n_codes, int32_t num_db, int64_t* d_out, void* stream);

/* Workspace bound used to size query chunks (default 8 GiB), and the bytes
 * one query needs -- the unit of the serving loop's HBM admission. */
int ck_kmer_set_workspace_budget(int64_t bytes);
int64_t ck_kmer_query_bytes(int32_t num_db, int32_t max_db_len, int32_t max_qlen, int32_t k,
                            int32_t diag_bin_width, int32_t max_cands, int32_t phase_a_topk);

/* Jaccard estimate over sorted bottom-s sketches for each candidate pair
 * (clustkit/pairwise.py:1564-1612, CUDA twin 1664-1712).  Device pointers:
 * pairs int32[2M], sketches uint64[N*s] row-major, out float32[M]. */
int ck_jaccard_device(const int32_t* d_pairs, int64_t M, const uint64_t* d_sketches,
                      int32_t sketch_size, float* d_out_sims, void* stream);

/* Host-pointer Jaccard (synchronous), n_seqs rows of sketch_size values. */
int ck_jaccard_host(const int32_t* h_pairs, int64_t M, const uint64_t* h_sketches,
                    int64_t n_seqs, int32_t sketch_size, float* h_out_sims);

/* Bottom-s MinHash sketches of a sequence table in HBM
 * (clustkit/sketch.py:17-80 and 137-183; CUDA twin sketch.py:193-281):
 * d_out_sketches uint64[n_seqs * sketch_size] row-major, each row the
 * sketch_size smallest fmix64(code ^ seed) over the sequence's k-mer windows,
 * ascending, padded with 2^64-1; all zeros for a sequence shorter than k.
 * alphabet_size is the base of the k-mer code (20 protein, 4 nucleotide).
 * sketch_size <= 1024.  Asynchronous on `stream`. */
int ck_sketch_device(const uint8_t* d_flat, const int64_t* d_offsets, const int32_t* d_lengths,
                     int64_t n_seqs, int32_t k, int32_t sketch_size, int32_t alphabet_size,
                     uint64_t seed, uint64_t* d_out_sketches, void* stream);

/* LSH candidate pairs (clustkit/lsh.py:93-178, hashing lsh.py:20-29, pair
 * extraction lsh.py:52-90): table t hashes the sketch positions
 * h_band_indices[t*num_bands ..] with seed h_seeds[t] (host arrays, drawn by
 * the caller as lsh.py:133-137 does); members of a bucket in ascending
 * sequence order, the first max_bucket of them (reference: 1000) paired.
 * Synchronous: *out_n_pairs = number of distinct pairs (i < j); the pairs
 * stay in the stream's workspace until the next call on that stream and are
 * written by ck_lsh_fetch_pairs_device. */
int ck_lsh_candidates_device(const uint64_t* d_sketches, int64_t n_seqs, int32_t sketch_size,
                             const int32_t* h_band_indices, const int64_t* h_seeds,
                             int32_t num_tables, int32_t num_bands, int32_t max_bucket,
                             int64_t* out_n_pairs, void* stream);

/* Writes the pairs of the last ck_lsh_candidates_device call on `stream`
 * as int32[n_pairs * 2] (i, j), ascending by i * n_seqs + j; n_pairs must be
 * the count that call returned.  Asynchronous on `stream`. */
int ck_lsh_fetch_pairs_device(int32_t* d_out_pairs, int64_t n_pairs, void* stream);

/* Rank of ck_search_topk_device (kmer_index.py:2310-2322 with
 * search.py:764): identity (no score, e.g. global alignment with +2/-1),
 * the E-value-style rank of SW scores, or the alignment score itself
 * (SW without evalue_normalize, global alignment with a table). */
#define CK_RANK_IDENTITY 0
#define CK_RANK_EVALUE   1
#define CK_RANK_SCORE    2

/* Top-k hits per query (clustkit/search.py:727-800, rank score of
 * clustkit/kmer_index.py:2318-2322).  Device pointers; query q has candidate
 * slots 0 .. d_counts[q]-1 of d_targets[q*max_cands ..] with identity
 * d_sims, score d_scores and mask d_mask at the same index; a slot enters
 * when its mask is set.  Rank (rank_mode): CK_RANK_IDENTITY the identity,
 * CK_RANK_SCORE fl32(score), CK_RANK_EVALUE
 * fl32(fl32(0.267f * score) - d_logtab[qlen * tlen]) (qlen from d_qlens[q],
 * tlen from d_tlens[target]; d_logtab[p] = float32 ln p as NumPy computes it,
 * logtab_len entries).  Order: rank descending, then slot ascending (the
 * stable lexsort by (query, -rank)); the first top_k go to
 * d_out_*[q*top_k ..], their number to d_out_n[q].  *d_status (zeroed by the
 * caller) gets bit 0 when a length product falls outside the table.
 * max_cands <= 2048.  Asynchronous on `stream`. */
int ck_search_topk_device(const int32_t* d_targets, const int32_t* d_counts,
                          const float* d_sims, const int32_t* d_scores, const uint8_t* d_mask,
                          const int32_t* d_qlens, const int32_t* d_tlens, const float* d_logtab,
                          int64_t logtab_len, int32_t nq, int32_t max_cands, int32_t top_k,
                          int32_t rank_mode, int32_t* d_out_target, float* d_out_identity,
                          float* d_out_rank, int32_t* d_out_n, int32_t* d_status, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CLUSTKIT_HIP_H */
