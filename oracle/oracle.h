/*
 * oracle.h -- CPU restatement of the BLASTP preliminary stage (test infrastructure).
 *
 * THIS IS THE PARITY ORACLE, NOT PRODUCT CODE.  Only tests/, the smoke check and the
 * cpu_baseline leg of bench.py may load it.  The product (the HIP library) never links it.
 *
 * What it restates, per (query, subject) pair, in the order NCBI BLAST+ 2.13.0 runs it
 * inside s_BlastSearchEngineOneContextMT (BLAZE_ncbi/blast_engine.c:653-853):
 *   - the word-3 / threshold-11 neighbourhood lookup table
 *       (BLAZE_ncbi/prelim_stage.cpp:135-155, 360-402; BLAZE_gpu/util.hpp:64-85)
 *   - the two-hit word finder with its per-diagonal last-hit / flag state
 *       (WordFinder call, blast_engine.c:731; NCBI aa_ungapped.c, not vendored)
 *   - the ungapped X-drop extension left then right
 *       (BLAZE_gpu/gpuBLAZE.cu:680-883 restates the same stop rules)
 *   - the gapped start (best 11-wide window), BLAZE_gpu/gpuBLAZE.cu:1159-1210
 *   - the score-only semi-gapped X-drop DP, left on reversed prefixes and right on
 *       suffixes (GetGappedScore call, blast_engine.c:767; BLAZE_gpu/gpuBLAZE.cu:356-677
 *       states the banded variant)
 *   - the init-hit ordering and HSP containment test, the common-endpoint purge
 *       (blast_engine.c:792) and the score sort (blast_engine.c:801).
 *
 * Coordinates are 0-based; every *_end is one past the last aligned residue.
 */
#ifndef BLZ_ORACLE_H
#define BLZ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_ALPHA 28
#define ORA_NWORDS (28 * 28 * 28)

typedef struct {
    int32_t word_threshold;   /* 11 */
    int32_t window;           /* 40: two-hit window */
    int32_t x_ungapped;       /* 16: raw ungapped X-drop */
    int32_t x_gapped;         /* 38: raw prelim gapped X-drop */
    int32_t gap_open;         /* 11 */
    int32_t gap_extend;       /* 1 */
    int32_t ungapped_cutoff;  /* raw score an ungapped HSP needs to become an init hit */
    int32_t gapped_cutoff;    /* raw score a gapped HSP needs to be saved */
} ora_params;

typedef struct {
    int32_t q_hit, s_hit;     /* start of the second word hit of the pair */
    int32_t q_start, s_start; /* ungapped segment */
    int32_t len;
    int32_t score;
} ora_init_hit;

typedef struct {
    int32_t score;
    int32_t q_off, q_end;     /* query range, end exclusive */
    int32_t s_off, s_end;     /* subject range, end exclusive */
    int32_t q_gs, s_gs;       /* gapped start the DP was anchored at */
} ora_hsp;

typedef struct {
    uint64_t word_hits;       /* lookup hits produced by the scan */
    uint64_t ungapped_ext;    /* two-hit extensions performed */
    uint64_t init_hits;       /* extensions that reached the ungapped cutoff */
    uint64_t gapped_ext;      /* semi-gapped extensions performed (left+right = 1) */
    uint64_t dp_cells;        /* DP cells evaluated inside the X-drop band */
    uint64_t hsps;            /* HSPs after purge */
} ora_stats;

/* Lookup table over the query: prefix[ORA_NWORDS+1] into entries (query offsets, ascending). */
typedef struct {
    uint32_t *prefix;
    int32_t *entries;
    int32_t qlen;
} ora_lookup;

void ora_default_params(ora_params *p);
const int32_t *ora_matrix(void);   /* 28x28 BLOSUM62, NCBIstdaa order, row-major */

/* Neighbourhood of word w: every u with score(w,u) >= T, plus w itself. Returns the count
 * written to out (out may be NULL to only count). */
int ora_neighbors(int w, int threshold, int32_t *out);

ora_lookup *ora_lookup_build(const uint8_t *query, int qlen, int threshold);
void ora_lookup_free(ora_lookup *lk);

/* Ungapped extension helpers (exposed for unit checks). */
int ora_extend_left(const uint8_t *q, const uint8_t *s, int q_off, int s_off, int dropoff,
                    int maxscore, int *length);
int ora_extend_right(const uint8_t *q, int qlen, const uint8_t *s, int slen, int q_off, int s_off,
                     int dropoff, int maxscore, int *length, int *s_last_off);

/* Two-hit word finder over one subject. Returns the number of init hits (all written if
 * cap allows; the return value may exceed cap). */
int ora_word_finder(const uint8_t *query, int qlen, const ora_lookup *lk, const uint8_t *subj,
                    int slen, const ora_params *p, ora_init_hit *out, int cap, ora_stats *st);

int ora_gapped_start(const uint8_t *q, const uint8_t *s, int q_start, int s_start, int len);

/* Score-only semi-gapped X-drop DP (NCBI Blast_SemiGappedAlign, score_only path).
 * reverse != 0 walks A[M-1..0] and B[N-1..0]. a_off/b_off receive the number of letters
 * of A and B consumed at the best cell. cells (may be NULL) accumulates evaluated cells. */
int ora_semi_gapped(const uint8_t *A, const uint8_t *B, int M, int N, int x_dropoff, int gap_open,
                    int gap_extend, int reverse, int *a_off, int *b_off, uint64_t *cells);

/* Whole prelim stage for one subject: returns the HSP count (written up to cap). */
int ora_prelim_subject(const uint8_t *query, int qlen, const ora_lookup *lk, const uint8_t *subj,
                       int slen, const ora_params *p, ora_hsp *out, int cap, ora_stats *st);

/* The gapped stage of ora_prelim_subject on a given list of init hits (NCBI
 * BLAST_GetGappedScore): init-hit order, containment, gapped start, both DPs, cutoff, purge,
 * sort.  Returns the HSP count (the first cap are written). */
int ora_gapped_stage(const uint8_t *query, int qlen, const uint8_t *subj, int slen,
                     const ora_init_hit *hits, int n, const ora_params *p, ora_hsp *out, int cap,
                     ora_stats *st);

/* Common-endpoint purge (NCBI Blast_HSPListPurgeHSPsWithCommonEndpoints, purge = TRUE): returns
 * the number kept, h[0..k) in query-end order.  Final order (NCBI ScoreCompareHSPs). */
int ora_purge(ora_hsp *h, int n);
void ora_sort_hsps(ora_hsp *h, int n);

/* ---- traceback stage (traceback.c) ---- */
enum { ORA_OP_SUB = 0, ORA_OP_GAP_Q = 1, ORA_OP_GAP_S = 2 };  /* aligned pair, subject letter alone, query letter alone */

typedef struct {
    int32_t score;
    int32_t q_off, q_end, s_off, s_end;  /* ends exclusive */
    int32_t q_gs, s_gs;                  /* where the alignment was anchored */
    int32_t num_ident, align_len, gaps, gap_opens;
    int32_t op_first, op_count;          /* its runs (op << 28 | length), left to right */
} ora_tb_hsp;

/* One direction of NCBI's gapped alignment with traceback (ALIGN_EX): score, best cell, and
 * the path's operations from the best cell back to the origin. */
int ora_align_ex(const uint8_t *A, const uint8_t *B, int M, int N, int x_dropoff, int gap_open, int gap_extend,
                 int reverse, int *a_off, int *b_off, uint8_t *ops, int *nops, uint64_t *cells);
/* Traceback of one subject's prelim HSPs (score order): containment against the HSPs traced so
 * far, alignment from each gapped start with the final X-drop, cutoff, purge and final order.
 * *nruns receives the runs written (negative: that many needed, runs_cap too small). */
int ora_traceback_subject(const uint8_t *q, int qlen, const uint8_t *s, int slen, const ora_hsp *prelim, int n,
                          int x_final, int gap_open, int gap_extend, int cutoff, ora_tb_hsp *out, int cap,
                          uint32_t *runs, int runs_cap, int *nruns, uint64_t *cells);

/* Prelim stage over a database of nsubj subjects (residues concatenated, offsets[nsubj+1]).
 * counts[i] receives subject i's HSP count; HSPs are written subject-major into out
 * (capacity cap). Returns the total HSP count, or -1 if cap was too small (counts are
 * still filled). nthreads <= 1 runs single-threaded. */
int64_t ora_prelim_db(const uint8_t *query, int qlen, const uint8_t *residues,
                      const uint64_t *offsets, int64_t nsubj, const ora_params *p, int nthreads,
                      int32_t *counts, ora_hsp *out, int64_t cap, ora_stats *st);

#ifdef __cplusplus
}
#endif
#endif
