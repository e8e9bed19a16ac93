/*
 * oracle.c -- CPU restatement of the BLASTP preliminary stage (test infrastructure only).
 *
 * Only tests/, the smoke check and bench.py's cpu_baseline leg load this file's library.
 * The product HIP library never links or calls it.
 *
 * Every function cites the reference code it follows.  NCBI BLAST+ 2.13.0 itself is a
 * third-party dependency that the reference downloads at build time (build.sh:29,51) and
 * does not vendor; where the semantics live only there (aa_ungapped.c, blast_gapalign.c,
 * blast_hits.c, blast_itree.c) this file restates NCBI's published algorithm and says so.
 */
#include "oracle.h"

#include <limits.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* BLOSUM62 in NCBIstdaa order (- A B C D E F G H I K L M N P Q R S T V W X Y Z U * O J).
 * Values are the reference's table, BLAZE_gpu/util.hpp:30-57 (row/column 0, the gap letter,
 * is -32768 there and -16384 in BLAZE_ncbi/prelim_stage.cpp:69-96; no synthetic sequence
 * contains letter 0, so either value gives the same results). */
static const int32_t kBlosum62[28][28] = {
#define G -32768
    {G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G, G},
#undef G
    {-32768, 4, -2, 0, -2, -1, -2, 0, -2, -1, -1, -1, -1, -2, -1, -1, -1, 1, 0, 0, -3, -1, -2, -1, 0, -4, -1, -1},
    {-32768, -2, 4, -3, 4, 1, -3, -1, 0, -3, 0, -4, -3, 4, -2, 0, -1, 0, -1, -3, -4, -1, -3, 0, -3, -4, -1, -3},
    {-32768, 0, -3, 9, -3, -4, -2, -3, -3, -1, -3, -1, -1, -3, -3, -3, -3, -1, -1, -1, -2, -1, -2, -3, 9, -4, -1, -1},
    {-32768, -2, 4, -3, 6, 2, -3, -1, -1, -3, -1, -4, -3, 1, -1, 0, -2, 0, -1, -3, -4, -1, -3, 1, -3, -4, -1, -3},
    {-32768, -1, 1, -4, 2, 5, -3, -2, 0, -3, 1, -3, -2, 0, -1, 2, 0, 0, -1, -2, -3, -1, -2, 4, -4, -4, -1, -3},
    {-32768, -2, -3, -2, -3, -3, 6, -3, -1, 0, -3, 0, 0, -3, -4, -3, -3, -2, -2, -1, 1, -1, 3, -3, -2, -4, -1, 0},
    {-32768, 0, -1, -3, -1, -2, -3, 6, -2, -4, -2, -4, -3, 0, -2, -2, -2, 0, -2, -3, -2, -1, -3, -2, -3, -4, -1, -4},
    {-32768, -2, 0, -3, -1, 0, -1, -2, 8, -3, -1, -3, -2, 1, -2, 0, 0, -1, -2, -3, -2, -1, 2, 0, -3, -4, -1, -3},
    {-32768, -1, -3, -1, -3, -3, 0, -4, -3, 4, -3, 2, 1, -3, -3, -3, -3, -2, -1, 3, -3, -1, -1, -3, -1, -4, -1, 3},
    {-32768, -1, 0, -3, -1, 1, -3, -2, -1, -3, 5, -2, -1, 0, -1, 1, 2, 0, -1, -2, -3, -1, -2, 1, -3, -4, -1, -3},
    {-32768, -1, -4, -1, -4, -3, 0, -4, -3, 2, -2, 4, 2, -3, -3, -2, -2, -2, -1, 1, -2, -1, -1, -3, -1, -4, -1, 3},
    {-32768, -1, -3, -1, -3, -2, 0, -3, -2, 1, -1, 2, 5, -2, -2, 0, -1, -1, -1, 1, -1, -1, -1, -1, -1, -4, -1, 2},
    {-32768, -2, 4, -3, 1, 0, -3, 0, 1, -3, 0, -3, -2, 6, -2, 0, 0, 1, 0, -3, -4, -1, -2, 0, -3, -4, -1, -3},
    {-32768, -1, -2, -3, -1, -1, -4, -2, -2, -3, -1, -3, -2, -2, 7, -1, -2, -1, -1, -2, -4, -1, -3, -1, -3, -4, -1, -3},
    {-32768, -1, 0, -3, 0, 2, -3, -2, 0, -3, 1, -2, 0, 0, -1, 5, 1, 0, -1, -2, -2, -1, -1, 4, -3, -4, -1, -2},
    {-32768, -1, -1, -3, -2, 0, -3, -2, 0, -3, 2, -2, -1, 0, -2, 1, 5, -1, -1, -3, -3, -1, -2, 0, -3, -4, -1, -2},
    {-32768, 1, 0, -1, 0, 0, -2, 0, -1, -2, 0, -2, -1, 1, -1, 0, -1, 4, 1, -2, -3, -1, -2, 0, -1, -4, -1, -2},
    {-32768, 0, -1, -1, -1, -1, -2, -2, -2, -1, -1, -1, -1, 0, -1, -1, -1, 1, 5, 0, -2, -1, -2, -1, -1, -4, -1, -1},
    {-32768, 0, -3, -1, -3, -2, -1, -3, -3, 3, -2, 1, 1, -3, -2, -2, -3, -2, 0, 4, -3, -1, -1, -2, -1, -4, -1, 2},
    {-32768, -3, -4, -2, -4, -3, 1, -2, -2, -3, -3, -2, -1, -4, -4, -2, -3, -3, -2, -3, 11, -1, 2, -2, -2, -4, -1, -2},
    {-32768, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -4, -1, -1},
    {-32768, -2, -3, -2, -3, -2, 3, -3, 2, -1, -2, -1, -1, -2, -3, -1, -2, -2, -2, -1, 2, -1, 7, -2, -2, -4, -1, -1},
    {-32768, -1, 0, -3, 1, 4, -3, -2, 0, -3, 1, -3, -1, 0, -1, 4, 0, 0, -1, -2, -2, -1, -2, 4, -3, -4, -1, -3},
    {-32768, 0, -3, 9, -3, -4, -2, -3, -3, -1, -3, -1, -1, -3, -3, -3, -3, -1, -1, -1, -2, -1, -2, -3, 9, -4, -1, -1},
    {-32768, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, -4, 1, -4, -4},
    {-32768, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -4, -1, -1},
    {-32768, -1, -3, -1, -3, -3, 0, -4, -3, 3, -3, 3, 2, -3, -3, -2, -2, -2, -1, 2, -2, -1, -1, -3, -1, -4, -1, 3},
};

#define SC(a, b) (kBlosum62[(a)][(b)])
#define MININT_DP (INT_MIN / 2) /* NCBI's MININT for the DP, BLAZE_gpu/gpuBLAZE.cu:75 */
#define WORD 3
#define HSP_MAX_WINDOW 11       /* BLAZE_gpu/gpuBLAZE.cu:64 */

void ora_default_params(ora_params *p)
{
    /* Defaults named by the reference: threshold 11 (prelim_stage.cpp:144), window 40 and
     * ungapped X 16 (gpuBLAZE.cu:59-61), gapped X 38 (gpuBLAZE.cu:37), gap 11/1 (:67-72).
     * The cutoffs are runtime inputs in the reference (gpu_cpu_common.h:43); 41 is the
     * ungapped gap trigger the reference hard-codes (gpuBLAZE.cu:63) and 67 its gapped one
     * (gpuBLAZE.cu:38). */
    p->word_threshold = 11;
    p->window = 40;
    p->x_ungapped = 16;
    p->x_gapped = 38;
    p->gap_open = 11;
    p->gap_extend = 1;
    p->ungapped_cutoff = 41;
    p->gapped_cutoff = 67;
}

const int32_t *ora_matrix(void) { return &kBlosum62[0][0]; }

/* ---------------------------------------------------------------------------------------
 * Neighbourhood and lookup table.
 * Follows BLAZE_ncbi/prelim_stage.cpp:135-155 (every word u with sum of BLOSUM62 >= 11,
 * self excluded) and :367-386 (the query word itself is added for its own position, so an
 * exact word always hits even when its self-score is below 11). Word index a*784+b*28+c
 * (gpuBLAZE.cu:991).  Entries per word are in ascending query offset, the order in which
 * NCBI's lookup indexing appends them (outer loop over query offsets).
 * ------------------------------------------------------------------------------------- */
int ora_neighbors(int w, int threshold, int32_t *out)
{
    int x = w / 784, y = (w / 28) % 28, z = w % 28;
    int rowmax_y = INT_MIN, rowmax_z = INT_MIN;
    for (int i = 0; i < 28; i++) {
        if (SC(y, i) > rowmax_y) rowmax_y = SC(y, i);
        if (SC(z, i) > rowmax_z) rowmax_z = SC(z, i);
    }
    int n = 0;
    /* the pruning below only skips branches whose best completion is below the threshold,
     * so the enumeration is exactly {u : score(w,u) >= threshold} */
    for (int a = 0; a < 28; a++) {
        int sa = SC(x, a);
        if (sa + rowmax_y + rowmax_z < threshold) continue;
        for (int b = 0; b < 28; b++) {
            int sb = sa + SC(y, b);
            if (sb + rowmax_z < threshold) continue;
            for (int c = 0; c < 28; c++) {
                if (sb + SC(z, c) >= threshold) {
                    if (out) out[n] = a * 784 + b * 28 + c;
                    n++;
                }
            }
        }
    }
    if (SC(x, x) + SC(y, y) + SC(z, z) < threshold) { /* exact word added on its own */
        if (out) out[n] = w;
        n++;
    }
    return n;
}

ora_lookup *ora_lookup_build(const uint8_t *query, int qlen, int threshold)
{
    ora_lookup *lk = (ora_lookup *)calloc(1, sizeof(ora_lookup));
    lk->qlen = qlen;
    lk->prefix = (uint32_t *)calloc(ORA_NWORDS + 1, sizeof(uint32_t));
    int nw = qlen >= WORD ? qlen - WORD + 1 : 0;
    /* neighbourhoods of the distinct query words, computed once each */
    int32_t **nb = (int32_t **)calloc(ORA_NWORDS, sizeof(int32_t *));
    int *nbn = (int *)calloc(ORA_NWORDS, sizeof(int));
    int32_t *scratch = (int32_t *)malloc(sizeof(int32_t) * (ORA_NWORDS + 1));
    for (int q = 0; q < nw; q++) {
        int w = query[q] * 784 + query[q + 1] * 28 + query[q + 2];
        if (!nb[w]) {
            int n = ora_neighbors(w, threshold, scratch);
            nb[w] = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
            memcpy(nb[w], scratch, sizeof(int32_t) * n);
            nbn[w] = n;
        }
        for (int k = 0; k < nbn[w]; k++) lk->prefix[nb[w][k] + 1]++;
    }
    for (int u = 0; u < ORA_NWORDS; u++) lk->prefix[u + 1] += lk->prefix[u];
    uint32_t total = lk->prefix[ORA_NWORDS];
    lk->entries = (int32_t *)malloc(sizeof(int32_t) * (total ? total : 1));
    uint32_t *cursor = (uint32_t *)malloc(sizeof(uint32_t) * ORA_NWORDS);
    memcpy(cursor, lk->prefix, sizeof(uint32_t) * ORA_NWORDS);
    for (int q = 0; q < nw; q++) { /* ascending q: each word's list comes out sorted */
        int w = query[q] * 784 + query[q + 1] * 28 + query[q + 2];
        for (int k = 0; k < nbn[w]; k++) lk->entries[cursor[nb[w][k]]++] = q;
    }
    for (int u = 0; u < ORA_NWORDS; u++) free(nb[u]);
    free(nb);
    free(nbn);
    free(scratch);
    free(cursor);
    return lk;
}

void ora_lookup_free(ora_lookup *lk)
{
    if (!lk) return;
    free(lk->prefix);
    free(lk->entries);
    free(lk);
}

/* ---------------------------------------------------------------------------------------
 * Ungapped X-drop extension.  NCBI aa_ungapped.c s_BlastAaExtendLeft/Right (not vendored);
 * the reference restates the same stop rules in BLAZE_gpu/gpuBLAZE.cu:680-778 (left: stop
 * once the drop from the best reaches the X-drop) and :780-883 (right: also stop once the
 * running score is <= 0).  The left walk starts at (q_off, s_off) inclusive and walks to
 * offset 0; *length is the number of letters left of and including the start that the best
 * prefix covers.  The right walk starts at (q_off, s_off) inclusive; *length is one past the
 * best position and *s_last_off the subject offset where the walk stopped.
 * ------------------------------------------------------------------------------------- */
int ora_extend_left(const uint8_t *q, const uint8_t *s, int q_off, int s_off, int dropoff,
                    int maxscore, int *length)
{
    int n = s_off < q_off ? s_off : q_off;
    int best_i = n + 1;
    int score = maxscore;
    for (int i = n; i >= 0; i--) {
        score += SC(q[q_off - n + i], s[s_off - n + i]);
        if (score > maxscore) {
            maxscore = score;
            best_i = i;
        }
        if (maxscore - score >= dropoff) break;
    }
    *length = n - best_i + 1;
    return maxscore;
}

int ora_extend_right(const uint8_t *q, int qlen, const uint8_t *s, int slen, int q_off, int s_off,
                     int dropoff, int maxscore, int *length, int *s_last_off)
{
    int n = slen - s_off < qlen - q_off ? slen - s_off : qlen - q_off;
    int best_i = -1;
    int score = maxscore;
    int i;
    for (i = 0; i < n; i++) {
        score += SC(q[q_off + i], s[s_off + i]);
        if (score > maxscore) {
            maxscore = score;
            best_i = i;
        }
        if (score <= 0 || maxscore - score >= dropoff) break;
    }
    *length = best_i + 1;
    *s_last_off = s_off + i;
    return maxscore;
}

/* Two-hit extension (NCBI s_BlastAaExtendTwoHit; the reference's restatement is
 * gpuBLAZE.cu:1116-1154).  The second hit's word is scanned for the prefix with the best
 * positive score; the left walk starts one before the end of that prefix; the right walk is
 * only made when the left walk reached the first hit (gpuBLAZE.cu:1143). */
static int two_hit_extend(const uint8_t *q, int qlen, const uint8_t *s, int slen, int s_left_off,
                          int s_right_off, int q_right_off, int dropoff, int *hsp_q, int *hsp_s,
                          int *hsp_len, int *right_extend, int *s_last_off)
{
    int left_d = 0, right_d = 0, left_score = 0, right_score = 0, score = 0;
    for (int i = 0; i < WORD; i++) {
        score += SC(q[q_right_off + i], s[s_right_off + i]);
        if (score > left_score) {
            left_score = score;
            right_d = i + 1;
        }
    }
    q_right_off += right_d;
    s_right_off += right_d;
    right_d = 0;
    *right_extend = 0;
    *s_last_off = s_right_off;
    left_score = ora_extend_left(q, s, q_right_off - 1, s_right_off - 1, dropoff, 0, &left_d);
    if (left_d >= s_right_off - s_left_off) {
        *right_extend = 1;
        right_score = ora_extend_right(q, qlen, s, slen, q_right_off, s_right_off, dropoff,
                                       left_score, &right_d, s_last_off);
    }
    *hsp_q = q_right_off - left_d;
    *hsp_s = s_right_off - left_d;
    *hsp_len = left_d + right_d;
    return left_score > right_score ? left_score : right_score;
}

/* ---------------------------------------------------------------------------------------
 * Two-hit word finder (NCBI aa_ungapped.c s_BlastAaWordFinder_TwoHit, called through
 * aux_struct->WordFinder at blast_engine.c:731).  Subject words are visited in ascending
 * subject offset, and each word's query offsets in ascending order.  Per diagonal q-s the
 * state is the last hit and a flag that is set after a right extension:
 *   flag set:   hits before the end of that extension are dropped; the first hit at or
 *               past it becomes the new last hit and clears the flag;
 *   flag clear: a hit >= window past the last hit becomes the new last hit; a hit that
 *               overlaps the last hit (distance < word size) is dropped; otherwise the pair
 *               is extended, and the last hit becomes either the end of the right
 *               extension minus (word size - 1) with the flag set, or this hit.
 * The reference's GPU filter drops the flag and always moves the last hit to the current
 * hit (gpuBLAZE.cu:1023-1049); the oracle keeps NCBI's rule.
 * ------------------------------------------------------------------------------------- */
int ora_word_finder(const uint8_t *query, int qlen, const ora_lookup *lk, const uint8_t *subj,
                    int slen, const ora_params *p, ora_init_hit *out, int cap, ora_stats *st)
{
    if (slen < WORD || qlen < WORD) return 0;
    int ndiag = qlen + slen;
    int32_t *last = (int32_t *)malloc(sizeof(int32_t) * ndiag);
    uint8_t *flag = (uint8_t *)calloc(ndiag, 1);
    for (int i = 0; i < ndiag; i++) last[i] = INT_MIN / 2;
    int count = 0;
    uint64_t hits = 0, exts = 0;
    for (int s = 0; s + WORD <= slen; s++) {
        int w = subj[s] * 784 + subj[s + 1] * 28 + subj[s + 2];
        for (uint32_t k = lk->prefix[w]; k < lk->prefix[w + 1]; k++) {
            int q = lk->entries[k];
            int d = q - s + slen - 1;
            hits++;
            if (flag[d]) {
                if (s < last[d]) continue;
                last[d] = s;
                flag[d] = 0;
                continue;
            }
            int diff = s - last[d];
            if (diff >= p->window) {
                last[d] = s;
                continue;
            }
            if (diff < WORD) continue;
            int hq, hs, hl, rext, slast;
            int score = two_hit_extend(query, qlen, subj, slen, last[d] + WORD, s, q, p->x_ungapped,
                                       &hq, &hs, &hl, &rext, &slast);
            exts++;
            if (score >= p->ungapped_cutoff) {
                if (count < cap) {
                    ora_init_hit *h = &out[count];
                    h->q_hit = q;
                    h->s_hit = s;
                    h->q_start = hq;
                    h->s_start = hs;
                    h->len = hl;
                    h->score = score;
                }
                count++;
            }
            if (rext) {
                flag[d] = 1;
                last[d] = slast - (WORD - 1);
            } else {
                last[d] = s;
            }
        }
    }
    free(last);
    free(flag);
    if (st) {
        st->word_hits += hits;
        st->ungapped_ext += exts;
        st->init_hits += (uint64_t)count;
    }
    return count;
}

/* Gapped start: the middle of a short ungapped HSP, else the centre of its best 11-letter
 * window when that window scores above 0, else the HSP start.  NCBI
 * BlastGetStartForGappedAlignment; restated by the reference at gpuBLAZE.cu:1166-1205. */
int ora_gapped_start(const uint8_t *q, const uint8_t *s, int q_start, int s_start, int len)
{
    if (len <= HSP_MAX_WINDOW) return q_start + len / 2;
    int score = 0;
    for (int i = 0; i < HSP_MAX_WINDOW; i++) score += SC(q[q_start + i], s[s_start + i]);
    int max_score = score;
    int max_offset = q_start + HSP_MAX_WINDOW - 1;
    for (int i = q_start + HSP_MAX_WINDOW; i < q_start + len; i++) {
        int k = i - q_start;
        score -= SC(q[i - HSP_MAX_WINDOW], s[s_start + k - HSP_MAX_WINDOW]);
        score += SC(q[i], s[s_start + k]);
        if (score > max_score) {
            max_score = score;
            max_offset = i;
        }
    }
    if (max_score > 0) return max_offset - HSP_MAX_WINDOW / 2;
    return q_start;
}

/* ---------------------------------------------------------------------------------------
 * Score-only semi-gapped X-drop DP, NCBI blast_gapalign.c Blast_SemiGappedAlign (score_only
 * path; not vendored).  The reference states the same banded X-drop in
 * BLAZE_gpu/gpuBLAZE.cu:356-515 (left) and :517-677 (right).  Rows walk A, the band walks B:
 *   row 0 is a run of gaps in A while it stays within the X-drop;
 *   a cell failing the X-drop is skipped by later rows (the first column of the band moves
 *   right when the failing cell is the band's first; otherwise its score becomes MININT);
 *   when a row's last cell survives, the band is extended by the row's gap run while that
 *   stays within the X-drop, plus one MININT cell; otherwise the band is cut after the
 *   row's last surviving cell.
 * Stops when no cell of a row survives.  Returns the best score; a_off/b_off are the
 * letters of A/B consumed at the first cell (row-major) that reached it.
 * ------------------------------------------------------------------------------------- */
int ora_semi_gapped(const uint8_t *A, const uint8_t *B, int M, int N, int x_dropoff, int gap_open,
                    int gap_extend, int reverse, int *a_off, int *b_off, uint64_t *cells)
{
    int gap_open_extend = gap_open + gap_extend;
    *a_off = 0;
    *b_off = 0;
    if (x_dropoff < gap_open_extend) x_dropoff = gap_open_extend;
    if (N <= 0 || M <= 0) return 0;

    int32_t *best = (int32_t *)malloc(sizeof(int32_t) * (N + 2));
    int32_t *best_gap = (int32_t *)malloc(sizeof(int32_t) * (N + 2));
    int score = -gap_open_extend;
    best[0] = 0;
    best_gap[0] = -gap_open_extend;
    int i;
    for (i = 1; i <= N; i++) {
        if (score < -x_dropoff) break;
        best[i] = score;
        best_gap[i] = score - gap_open_extend;
        score -= gap_extend;
    }
    int b_size = i;
    int best_score = 0;
    int first_b_index = 0;
    uint64_t ncells = 0;

    for (int a_index = 1; a_index <= M; a_index++) {
        const int32_t *mrow = &kBlosum62[reverse ? A[M - a_index] : A[a_index]][0];
        int sc = MININT_DP;
        int score_gap_row = MININT_DP;
        int last_b_index = first_b_index;
        for (int b_index = first_b_index; b_index < b_size; b_index++) {
            int score_gap_col = best_gap[b_index];
            /* diagonal score of the next column; the last column (b_index == N) has none and
             * its value is never used, so no letter past the sequence is read */
            int next_score = b_index < N ? best[b_index] + mrow[reverse ? B[N - 1 - b_index] : B[b_index + 1]]
                                         : MININT_DP;
            ncells++;
            if (sc < score_gap_col) sc = score_gap_col;
            if (sc < score_gap_row) sc = score_gap_row;
            if (best_score - sc > x_dropoff) {
                if (b_index == first_b_index)
                    first_b_index++;
                else
                    best[b_index] = MININT_DP;
            } else {
                last_b_index = b_index;
                if (sc > best_score) {
                    best_score = sc;
                    *a_off = a_index;
                    *b_off = b_index;
                }
                score_gap_row -= gap_extend;
                score_gap_col -= gap_extend;
                best_gap[b_index] = (sc - gap_open_extend > score_gap_col) ? sc - gap_open_extend
                                                                           : score_gap_col;
                score_gap_row = (sc - gap_open_extend > score_gap_row) ? sc - gap_open_extend
                                                                       : score_gap_row;
                best[b_index] = sc;
            }
            sc = next_score;
        }
        if (first_b_index == b_size) break;
        if (last_b_index < b_size - 1) {
            b_size = last_b_index + 1;
        } else {
            while (score_gap_row >= best_score - x_dropoff && b_size <= N) {
                best[b_size] = score_gap_row;
                best_gap[b_size] = score_gap_row - gap_open_extend;
                score_gap_row -= gap_extend;
                b_size++;
            }
        }
        if (b_size <= N) {
            best[b_size] = MININT_DP;
            best_gap[b_size] = MININT_DP;
            b_size++;
        }
    }
    free(best);
    free(best_gap);
    if (cells) *cells += ncells;
    return best_score;
}

/* ---------------------------------------------------------------------------------------
 * Gapped stage for one subject (NCBI blast_gapalign.c BLAST_GetGappedScore, called at
 * blast_engine.c:767; not vendored).
 *  1. init hits sorted by ungapped score desc, then subject start asc, length desc, query
 *     start asc (NCBI score_compare_match);
 *  2. an init hit whose ungapped segment (end inclusive) has both end points inside the box
 *     of an HSP already saved with a score >= its ungapped score is skipped (NCBI
 *     BlastIntervalTreeContainsHSP / s_HSPIsContained, min_diag_separation 0 for blastp);
 *  3. otherwise the gapped start is taken on the hit's diagonal, the left DP runs over the
 *     reversed prefixes that end at the start (start included) and the right DP over the
 *     suffixes after it (BLAZE restates this split at gpuBLAZE.cu:1315-1350);
 *  4. the HSP is saved when left+right reaches the gapped cutoff.
 * ------------------------------------------------------------------------------------- */
static int cmp_init_hit(const void *va, const void *vb)
{
    const ora_init_hit *a = (const ora_init_hit *)va, *b = (const ora_init_hit *)vb;
    if (a->score != b->score) return a->score > b->score ? -1 : 1;
    if (a->s_start != b->s_start) return a->s_start < b->s_start ? -1 : 1;
    if (a->len != b->len) return a->len > b->len ? -1 : 1;
    if (a->q_start != b->q_start) return a->q_start < b->q_start ? -1 : 1;
    return 0;
}

static int contained_in(const ora_hsp *t, int q0, int q1, int s0, int s1, int score)
{
    return score <= t->score && t->q_off <= q0 && q0 <= t->q_end && t->s_off <= s0 &&
           s0 <= t->s_end && t->q_off <= q1 && q1 <= t->q_end && t->s_off <= s1 && s1 <= t->s_end;
}

/* Purge of HSPs with a common start or a common end (NCBI blast_hits.c
 * Blast_HSPListPurgeHSPsWithCommonEndpoints with purge = TRUE, blast_engine.c:792): within a
 * group sharing (query start, subject start) the first HSP in the order score desc, query end
 * desc, subject end desc survives; then within a group sharing (query end, subject end) the
 * first in the order score desc, query start desc, subject start desc survives.  The gapped
 * start breaks any remaining tie (ascending) so the survivor is fixed. */
static int cmp_start(const void *va, const void *vb)
{
    const ora_hsp *a = (const ora_hsp *)va, *b = (const ora_hsp *)vb;
    if (a->q_off != b->q_off) return a->q_off < b->q_off ? -1 : 1;
    if (a->s_off != b->s_off) return a->s_off < b->s_off ? -1 : 1;
    if (a->score != b->score) return a->score > b->score ? -1 : 1;
    if (a->q_end != b->q_end) return a->q_end > b->q_end ? -1 : 1;
    if (a->s_end != b->s_end) return a->s_end > b->s_end ? -1 : 1;
    if (a->q_gs != b->q_gs) return a->q_gs < b->q_gs ? -1 : 1;
    if (a->s_gs != b->s_gs) return a->s_gs < b->s_gs ? -1 : 1;
    return 0;
}

static int cmp_end(const void *va, const void *vb)
{
    const ora_hsp *a = (const ora_hsp *)va, *b = (const ora_hsp *)vb;
    if (a->q_end != b->q_end) return a->q_end < b->q_end ? -1 : 1;
    if (a->s_end != b->s_end) return a->s_end < b->s_end ? -1 : 1;
    if (a->score != b->score) return a->score > b->score ? -1 : 1;
    if (a->q_off != b->q_off) return a->q_off > b->q_off ? -1 : 1;
    if (a->s_off != b->s_off) return a->s_off > b->s_off ? -1 : 1;
    if (a->q_gs != b->q_gs) return a->q_gs < b->q_gs ? -1 : 1;
    if (a->s_gs != b->s_gs) return a->s_gs < b->s_gs ? -1 : 1;
    return 0;
}

/* Final order (NCBI ScoreCompareHSPs via Blast_HSPListSortByScore, blast_engine.c:801):
 * score desc, subject start asc, subject end desc, query start asc, query end desc. */
static int cmp_score(const void *va, const void *vb)
{
    const ora_hsp *a = (const ora_hsp *)va, *b = (const ora_hsp *)vb;
    if (a->score != b->score) return a->score > b->score ? -1 : 1;
    if (a->s_off != b->s_off) return a->s_off < b->s_off ? -1 : 1;
    if (a->s_end != b->s_end) return a->s_end > b->s_end ? -1 : 1;
    if (a->q_off != b->q_off) return a->q_off < b->q_off ? -1 : 1;
    if (a->q_end != b->q_end) return a->q_end > b->q_end ? -1 : 1;
    if (a->q_gs != b->q_gs) return a->q_gs < b->q_gs ? -1 : 1;
    if (a->s_gs != b->s_gs) return a->s_gs < b->s_gs ? -1 : 1;
    return 0;
}

static int purge_common_endpoints(ora_hsp *h, int n)
{
    if (n <= 1) return n;
    qsort(h, n, sizeof(ora_hsp), cmp_start);
    int m = 0;
    for (int i = 0; i < n; i++) {
        if (m > 0 && h[m - 1].q_off == h[i].q_off && h[m - 1].s_off == h[i].s_off) continue;
        h[m++] = h[i];
    }
    n = m;
    qsort(h, n, sizeof(ora_hsp), cmp_end);
    m = 0;
    for (int i = 0; i < n; i++) {
        if (m > 0 && h[m - 1].q_end == h[i].q_end && h[m - 1].s_end == h[i].s_end) continue;
        h[m++] = h[i];
    }
    return m;
}

int ora_purge(ora_hsp *h, int n) { return purge_common_endpoints(h, n); }

void ora_sort_hsps(ora_hsp *h, int n)
{
    if (n > 1) qsort(h, n, sizeof(ora_hsp), cmp_score);
}

int ora_prelim_subject(const uint8_t *query, int qlen, const ora_lookup *lk, const uint8_t *subj,
                       int slen, const ora_params *p, ora_hsp *out, int cap, ora_stats *st)
{
    int icap = 256;
    ora_init_hit *ih = (ora_init_hit *)malloc(sizeof(ora_init_hit) * icap);
    int n = ora_word_finder(query, qlen, lk, subj, slen, p, ih, icap, st);
    if (n > icap) {
        free(ih);
        icap = n;
        ih = (ora_init_hit *)malloc(sizeof(ora_init_hit) * icap);
        ora_stats dummy = {0};
        ora_word_finder(query, qlen, lk, subj, slen, p, ih, icap, &dummy);
    }
    int ns = ora_gapped_stage(query, qlen, subj, slen, ih, n, p, out, cap, st);
    free(ih);
    return ns;
}

int ora_gapped_stage(const uint8_t *query, int qlen, const uint8_t *subj, int slen,
                     const ora_init_hit *hits, int n, const ora_params *p, ora_hsp *out, int cap,
                     ora_stats *st)
{
    if (n <= 0) return 0;
    ora_init_hit *ih = (ora_init_hit *)malloc(sizeof(ora_init_hit) * n);
    memcpy(ih, hits, sizeof(ora_init_hit) * n);
    qsort(ih, n, sizeof(ora_init_hit), cmp_init_hit);
    ora_hsp *saved = (ora_hsp *)malloc(sizeof(ora_hsp) * n);
    int ns = 0;
    uint64_t gext = 0, cells = 0;
    for (int k = 0; k < n; k++) {
        const ora_init_hit *h = &ih[k];
        int q0 = h->q_start, q1 = h->q_start + h->len - 1;
        int s0 = h->s_start, s1 = h->s_start + h->len - 1;
        int skip = 0;
        for (int t = 0; t < ns && !skip; t++) skip = contained_in(&saved[t], q0, q1, s0, s1, h->score);
        if (skip) continue;
        int q_gs = ora_gapped_start(query, subj, h->q_start, h->s_start, h->len);
        int s_gs = h->s_start + (q_gs - h->q_start);
        int la, lb, ra, rb;
        int left = ora_semi_gapped(query, subj, q_gs + 1, s_gs + 1, p->x_gapped, p->gap_open,
                                   p->gap_extend, 1, &la, &lb, &cells);
        int right = ora_semi_gapped(query + q_gs, subj + s_gs, qlen - q_gs - 1, slen - s_gs - 1,
                                    p->x_gapped, p->gap_open, p->gap_extend, 0, &ra, &rb, &cells);
        gext++;
        int score = left + right;
        if (score >= p->gapped_cutoff) {
            ora_hsp *o = &saved[ns++];
            o->score = score;
            o->q_off = q_gs + 1 - la;
            o->q_end = q_gs + 1 + ra;
            o->s_off = s_gs + 1 - lb;
            o->s_end = s_gs + 1 + rb;
            o->q_gs = q_gs;
            o->s_gs = s_gs;
        }
    }
    free(ih);
    ns = purge_common_endpoints(saved, ns);
    qsort(saved, ns, sizeof(ora_hsp), cmp_score);
    for (int k = 0; k < ns && k < cap; k++) out[k] = saved[k];
    free(saved);
    if (st) {
        st->gapped_ext += gext;
        st->dp_cells += cells;
        st->hsps += (uint64_t)ns;
    }
    return ns;
}

/* ---------------------------------------------------------------------------------------
 * Database driver: subjects are independent, so threads take them from a shared counter
 * (the CPU-role loop of BLAST_PreliminarySearchEngineMT, blast_engine.c:2561-2893, visits
 * the same per-subject core).  Results are gathered per subject and written in subject
 * order.
 * ------------------------------------------------------------------------------------- */
typedef struct {
    const uint8_t *query;
    int qlen;
    const ora_lookup *lk;
    const uint8_t *residues;
    const uint64_t *offsets;
    int64_t nsubj;
    const ora_params *p;
    _Atomic int64_t next;
    ora_hsp **per;      /* per-subject HSP arrays (NULL if none) */
    int32_t *counts;
    ora_stats *tstats;  /* one per thread */
} db_job;

typedef struct {
    db_job *job;
    int tid;
} db_arg;

static void *db_worker(void *va)
{
    db_arg *a = (db_arg *)va;
    db_job *j = a->job;
    ora_stats *st = &j->tstats[a->tid];
    int cap = 64;
    ora_hsp *buf = (ora_hsp *)malloc(sizeof(ora_hsp) * cap);
    for (;;) {
        int64_t i0 = atomic_fetch_add(&j->next, 64);
        if (i0 >= j->nsubj) break;
        int64_t i1 = i0 + 64 < j->nsubj ? i0 + 64 : j->nsubj;
        for (int64_t i = i0; i < i1; i++) {
            const uint8_t *s = j->residues + j->offsets[i];
            int slen = (int)(j->offsets[i + 1] - j->offsets[i]);
            int n = ora_prelim_subject(j->query, j->qlen, j->lk, s, slen, j->p, buf, cap, st);
            if (n > cap) {
                free(buf);
                cap = n;
                buf = (ora_hsp *)malloc(sizeof(ora_hsp) * cap);
                ora_stats dummy = {0};
                ora_prelim_subject(j->query, j->qlen, j->lk, s, slen, j->p, buf, cap, &dummy);
            }
            j->counts[i] = n;
            if (n) {
                j->per[i] = (ora_hsp *)malloc(sizeof(ora_hsp) * n);
                memcpy(j->per[i], buf, sizeof(ora_hsp) * n);
            }
        }
    }
    free(buf);
    return NULL;
}

int64_t ora_prelim_db(const uint8_t *query, int qlen, const uint8_t *residues,
                      const uint64_t *offsets, int64_t nsubj, const ora_params *p, int nthreads,
                      int32_t *counts, ora_hsp *out, int64_t cap, ora_stats *st)
{
    ora_lookup *lk = ora_lookup_build(query, qlen, p->word_threshold);
    if (nthreads < 1) nthreads = 1;
    db_job job;
    job.query = query;
    job.qlen = qlen;
    job.lk = lk;
    job.residues = residues;
    job.offsets = offsets;
    job.nsubj = nsubj;
    job.p = p;
    atomic_init(&job.next, 0);
    job.per = (ora_hsp **)calloc(nsubj > 0 ? nsubj : 1, sizeof(ora_hsp *));
    job.counts = counts;
    job.tstats = (ora_stats *)calloc(nthreads, sizeof(ora_stats));
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    db_arg *args = (db_arg *)malloc(sizeof(db_arg) * nthreads);
    for (int t = 0; t < nthreads; t++) {
        args[t].job = &job;
        args[t].tid = t;
        if (t > 0) pthread_create(&th[t], NULL, db_worker, &args[t]);
    }
    db_worker(&args[0]);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
    int64_t total = 0;
    int fits = 1;
    for (int64_t i = 0; i < nsubj; i++) {
        int n = counts[i];
        if (n) {
            if (total + n <= cap && fits)
                memcpy(out + total, job.per[i], sizeof(ora_hsp) * n);
            else
                fits = 0;
            free(job.per[i]);
        }
        total += n;
    }
    if (st) {
        for (int t = 0; t < nthreads; t++) {
            st->word_hits += job.tstats[t].word_hits;
            st->ungapped_ext += job.tstats[t].ungapped_ext;
            st->init_hits += job.tstats[t].init_hits;
            st->gapped_ext += job.tstats[t].gapped_ext;
            st->dp_cells += job.tstats[t].dp_cells;
            st->hsps += job.tstats[t].hsps;
        }
    }
    free(job.per);
    free(job.tstats);
    free(th);
    free(args);
    ora_lookup_free(lk);
    return fits ? total : -1;
}
