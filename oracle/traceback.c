/*
 * traceback.c -- CPU restatement of the traceback stage (test infrastructure only).
 *
 * Only tests/ load this code (through liboracle.so), as the checker of the library's traceback
 * kernel.  The stage follows NCBI BLAST+ 2.13.0 (not vendored by the reference; it is the
 * unchanged stage downstream of the hook, the consumer of the lists BLAZE writes at
 * BLAZE_ncbi/blast_engine.c:2499): Blast_TracebackFromHSPList without composition-based
 * rescoring, its gapped alignment BLAST_GappedAlignmentWithTraceback and the dynamic program
 * ALIGN_EX (blast_gapalign.c), restated here from NCBI's published algorithm.
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

#define MININT_DP (INT_MIN / 2)

/* Traceback byte of a cell: the operation the cell's best score came from (low bits) and whether
 * its column and row gap scores extend an open gap rather than open one there. */
enum { T_SUB = 0, T_GAP_IN_A = 1, T_GAP_IN_B = 2, T_OP_MASK = 3, T_EXT_A = 4, T_EXT_B = 8 };

/* One direction of the gapped alignment with traceback.  The DP is Blast_SemiGappedAlign's
 * (ora_semi_gapped: rows walk A, the X-drop band walks B, the same updates and band rules) with a
 * traceback byte kept for every cell of every row: which of the diagonal, the column gap (B
 * alone: a gap in A) or the row gap (A alone: a gap in B) gave the cell's score -- the diagonal
 * unless a gap score is strictly greater, the row gap last -- and whether each gap score extends
 * its gap (the gap score minus the extension is at least the cell's score minus the open cost).
 * Row 0 is a gap run in A.  The path is then read back from the best cell (the first row-major
 * cell reaching the best score) to the origin: a cell reached through a gap continues that gap
 * while its extension bit is set.  ops receives one operation per alignment column in the order
 * the walk meets them (from the best cell back to the origin): ORA_OP_SUB, ORA_OP_GAP_Q (a B
 * letter against a gap in A) or ORA_OP_GAP_S (an A letter against a gap in B). */
int ora_align_ex(const uint8_t *A, const uint8_t *B, int M, int N, int x_dropoff, int gap_open, int gap_extend,
                 int reverse, int *a_off, int *b_off, uint8_t *ops, int *nops, uint64_t *cells)
{
    const int32_t *mat = ora_matrix();
    const int goe = gap_open + gap_extend;
    *a_off = 0;
    *b_off = 0;
    *nops = 0;
    if (x_dropoff < goe) x_dropoff = goe;
    if (N <= 0 || M <= 0) return 0;

    int32_t *best = (int32_t *)malloc(sizeof(int32_t) * (N + 2));
    int32_t *best_gap = (int32_t *)malloc(sizeof(int32_t) * (N + 2));
    /* traceback rows: row a holds cells [row_start[a], row_start[a] + row_len[a]) */
    uint8_t **row = (uint8_t **)calloc((size_t)M + 1, sizeof(uint8_t *));
    int32_t *row_start = (int32_t *)calloc((size_t)M + 1, sizeof(int32_t));

    int score = -goe;
    best[0] = 0;
    best_gap[0] = -goe;
    int i;
    for (i = 1; i <= N; i++) {
        if (score < -x_dropoff) break;
        best[i] = score;
        best_gap[i] = score - goe;
        score -= gap_extend;
    }
    int b_size = i;
    row[0] = (uint8_t *)malloc((size_t)b_size);
    for (int b = 0; b < b_size; b++) row[0][b] = T_GAP_IN_A;
    int best_score = 0, first_b_index = 0, rows = 0;
    uint64_t ncells = 0;
    for (int a_index = 1; a_index <= M; a_index++) {
        /* the row's cells: the live band plus the gap run that may be appended after it */
        const int extra = x_dropoff / gap_extend + 3;
        row_start[a_index] = first_b_index;
        row[a_index] = (uint8_t *)malloc((size_t)(b_size - first_b_index + extra + 1));
        uint8_t *tr = row[a_index] - first_b_index;
        rows = a_index;
        const int32_t *mrow = &mat[28 * (reverse ? A[M - a_index] : A[a_index])];
        int sc = MININT_DP, score_gap_row = MININT_DP;
        int last_b_index = first_b_index;
        for (int b_index = first_b_index; b_index < b_size; b_index++) {
            int score_gap_col = best_gap[b_index];
            const int next_score = b_index < N ? best[b_index] + mrow[reverse ? B[N - 1 - b_index] : B[b_index + 1]]
                                               : MININT_DP;
            uint8_t t = T_SUB;
            ncells++;
            if (sc < score_gap_col) {
                t = T_GAP_IN_B;
                sc = score_gap_col;
            }
            if (sc < score_gap_row) {
                t = T_GAP_IN_A;
                sc = score_gap_row;
            }
            if (best_score - sc > x_dropoff) {
                if (b_index == first_b_index)
                    first_b_index++;
                else
                    best[b_index] = MININT_DP;
            } else {
                if (sc > best_score) {
                    best_score = sc;
                    *a_off = a_index;
                    *b_off = b_index;
                }
                last_b_index = b_index;
                score_gap_row -= gap_extend;
                score_gap_col -= gap_extend;
                if (score_gap_col < sc - goe) {
                    best_gap[b_index] = sc - goe;
                } else {
                    best_gap[b_index] = score_gap_col;
                    t |= T_EXT_B;
                }
                if (score_gap_row < sc - goe)
                    score_gap_row = sc - goe;
                else
                    t |= T_EXT_A;
                best[b_index] = sc;
            }
            sc = next_score;
            tr[b_index] = t;
        }
        if (first_b_index == b_size) break;
        if (last_b_index < b_size - 1) {
            b_size = last_b_index + 1;
        } else {
            while (score_gap_row >= best_score - x_dropoff && b_size <= N) {
                best[b_size] = score_gap_row;
                best_gap[b_size] = score_gap_row - goe;
                score_gap_row -= gap_extend;
                tr[b_size] = T_GAP_IN_A;
                b_size++;
            }
        }
        if (b_size <= N) {
            best[b_size] = MININT_DP;
            best_gap[b_size] = MININT_DP;
            b_size++;
        }
    }
    /* the path, from the best cell back to the origin */
    int a = *a_off, b = *b_off, cur = T_SUB, n = 0;
    while (a > 0 || b > 0) {
        const uint8_t next = row[a][b - row_start[a]];
        if (cur == T_GAP_IN_A)
            cur = (next & T_EXT_A) ? T_GAP_IN_A : (next & T_OP_MASK);
        else if (cur == T_GAP_IN_B)
            cur = (next & T_EXT_B) ? T_GAP_IN_B : (next & T_OP_MASK);
        else
            cur = next & T_OP_MASK;
        if (cur == T_GAP_IN_A) {
            b--;
            ops[n++] = ORA_OP_GAP_Q;
        } else if (cur == T_GAP_IN_B) {
            a--;
            ops[n++] = ORA_OP_GAP_S;
        } else {
            a--;
            b--;
            ops[n++] = ORA_OP_SUB;
        }
    }
    *nops = n;
    for (int k = 0; k <= rows; k++) free(row[k]);
    free(row);
    free(row_start);
    free(best);
    free(best_gap);
    if (cells) *cells += ncells;
    return best_score;
}

static int contained_hsp(const ora_tb_hsp *t, const ora_hsp *h)
{
    /* NCBI s_HSPIsContained: score no higher, both end points inside the traced HSP's box */
    return h->score <= t->score && t->q_off <= h->q_off && h->q_off <= t->q_end && t->s_off <= h->s_off &&
           h->s_off <= t->s_end && t->q_off <= h->q_end && h->q_end <= t->q_end && t->s_off <= h->s_end &&
           h->s_end <= t->s_end;
}

/* Purge and final order of traced HSPs: the same rules as the prelim stage's (NCBI
 * Blast_HSPListPurgeHSPsWithCommonEndpoints, ScoreCompareHSPs), on their coordinates. */
static int tb_as_hsp_cmp(const ora_tb_hsp *x, ora_hsp *h)
{
    h->score = x->score;
    h->q_off = x->q_off;
    h->q_end = x->q_end;
    h->s_off = x->s_off;
    h->s_end = x->s_end;
    h->q_gs = x->q_gs;
    h->s_gs = x->s_gs;
    return 0;
}

int ora_traceback_subject(const uint8_t *q, int qlen, const uint8_t *s, int slen, const ora_hsp *prelim, int n,
                          int x_final, int gap_open, int gap_extend, int cutoff, ora_tb_hsp *out, int cap,
                          uint32_t *runs, int runs_cap, int *nruns, uint64_t *cells)
{
    const int32_t *mat = ora_matrix();
    ora_tb_hsp *tr = (ora_tb_hsp *)malloc(sizeof(ora_tb_hsp) * (n ? n : 1));
    uint8_t *opsL = (uint8_t *)malloc((size_t)qlen + slen + 8);
    uint8_t *opsR = (uint8_t *)malloc((size_t)qlen + slen + 8);
    uint8_t *cols = (uint8_t *)malloc((size_t)qlen + slen + 8);
    uint32_t **own = (uint32_t **)calloc(n ? n : 1, sizeof(uint32_t *));
    int *own_n = (int *)calloc(n ? n : 1, sizeof(int));
    int nt = 0;
    for (int k = 0; k < n; k++) {
        const ora_hsp *h = &prelim[k];
        int skip = 0;
        for (int t = 0; t < nt && !skip; t++) skip = contained_hsp(&tr[t], h);
        if (skip) continue;
        const int q_gs = h->q_gs, s_gs = h->s_gs;
        int la, lb, ra, rb, nl, nr;
        const int left = ora_align_ex(q, s, q_gs + 1, s_gs + 1, x_final, gap_open, gap_extend, 1, &la, &lb, opsL, &nl,
                                      cells);
        const int right = ora_align_ex(q + q_gs, s + s_gs, qlen - q_gs - 1, slen - s_gs - 1, x_final, gap_open,
                                       gap_extend, 0, &ra, &rb, opsR, &nr, cells);
        const int score = left + right;
        if (score < cutoff) continue;
        /* columns left to right: the left walk already runs in that order, the right one backwards */
        int nc = 0;
        for (int i = 0; i < nl; i++) cols[nc++] = opsL[i];
        for (int i = nr - 1; i >= 0; i--) cols[nc++] = opsR[i];
        ora_tb_hsp *o = &tr[nt];
        memset(o, 0, sizeof(*o));
        o->score = score;
        o->q_off = q_gs + 1 - la;
        o->q_end = q_gs + 1 + ra;
        o->s_off = s_gs + 1 - lb;
        o->s_end = s_gs + 1 + rb;
        o->q_gs = q_gs;
        o->s_gs = s_gs;
        o->align_len = nc;
        /* identities, gaps and the run-length form of the columns */
        own[nt] = (uint32_t *)malloc(sizeof(uint32_t) * (nc ? nc : 1));
        int qi = o->q_off, si = o->s_off, nrun = 0;
        for (int i = 0; i < nc; i++) {
            const uint8_t op = cols[i];
            if (op == ORA_OP_SUB) {
                o->num_ident += q[qi] == s[si];
                qi++;
                si++;
            } else {
                o->gaps++;
                if (i == 0 || cols[i - 1] != op) o->gap_opens++;
                if (op == ORA_OP_GAP_Q)
                    si++;
                else
                    qi++;
            }
            if (nrun > 0 && (own[nt][nrun - 1] >> 28) == op)
                own[nt][nrun - 1]++;
            else
                own[nt][nrun++] = ((uint32_t)op << 28) | 1u;
        }
        (void)mat;
        own_n[nt] = nrun;
        o->op_first = nt;  /* provisional: index of its own run list */
        nt++;
    }
    /* purge and order through the prelim comparators */
    ora_hsp *hs = (ora_hsp *)malloc(sizeof(ora_hsp) * (nt ? nt : 1));
    for (int t = 0; t < nt; t++) tb_as_hsp_cmp(&tr[t], &hs[t]);
    int kept = ora_purge(hs, nt);
    ora_sort_hsps(hs, kept);
    int total_runs = 0, fits = 1;
    for (int k = 0; k < kept; k++) {
        int t = 0;
        while (t < nt && !(tr[t].q_off == hs[k].q_off && tr[t].q_end == hs[k].q_end && tr[t].s_off == hs[k].s_off &&
                           tr[t].s_end == hs[k].s_end && tr[t].score == hs[k].score && tr[t].q_gs == hs[k].q_gs &&
                           tr[t].s_gs == hs[k].s_gs))
            t++;
        if (k < cap) {
            out[k] = tr[t];
            out[k].op_first = total_runs;
            out[k].op_count = own_n[t];
        }
        if (total_runs + own_n[t] <= runs_cap && k < cap)
            memcpy(runs + total_runs, own[t], sizeof(uint32_t) * own_n[t]);
        else
            fits = 0;
        total_runs += own_n[t];
    }
    *nruns = fits ? total_runs : -total_runs;
    for (int t = 0; t < nt; t++) free(own[t]);
    free(own);
    free(own_n);
    free(hs);
    free(tr);
    free(opsL);
    free(opsR);
    free(cols);
    return kept;
}
