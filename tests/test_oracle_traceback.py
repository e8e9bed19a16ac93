"""Pins the oracle's traceback stage (oracle/traceback.c).

ALIGN_EX's dynamic program is Blast_SemiGappedAlign's with a traceback byte per cell, so its score
and best cell must equal the (already pinned) score-only DP's for every X-drop; the path it reads
back must consume exactly the letters up to the best cell and re-score to the reported score
under the affine gap costs (a gap of k letters costs open + k * extend); with no X-drop it is an
optimal alignment (the Gotoh optimum of tests/test_oracle.py)."""
import numpy as np
import pytest

import oracle
from test_oracle import gotoh_anchored, mutate, rand_seq

M = oracle.matrix()


def rescore(A, B, ops, reverse, Mlen, Nlen, go=11, ge=1):
    """Score of the path (ops from the best cell back to the origin) and the letters it consumes."""
    a = b = 0
    score, prev = 0, None
    for op in ops[::-1]:  # origin to best cell
        if op == oracle.OP_SUB:
            a += 1
            b += 1
            ra = A[Mlen - a] if reverse else A[a]
            rb = B[Nlen - b] if reverse else B[b]
            score += int(M[ra, rb])
        else:
            if op == oracle.OP_GAP_Q:
                b += 1
            else:
                a += 1
            score -= ge + (go if op != prev else 0)
        prev = op
    return score, a, b


def indel_pair(rng, m):
    A = rand_seq(rng, m)
    B = mutate(rng, A, float(rng.uniform(0.3, 0.9)))
    for _ in range(int(rng.integers(0, 3))):
        if len(B) < 3:
            break
        k = int(rng.integers(1, len(B)))
        if rng.random() < 0.5:
            B = np.concatenate([B[:k], rand_seq(rng, int(rng.integers(1, 6))), B[k:]])
        else:
            B = np.concatenate([B[:k], B[k + int(rng.integers(1, 4)):]])
    return A, B


@pytest.mark.parametrize("reverse", [False, True])
def test_align_ex_equals_score_only_dp_and_path_rescores(reverse):
    rng = np.random.default_rng(21 + reverse)
    for trial in range(200):
        A, B = indel_pair(rng, int(rng.integers(4, 90)))
        x = int(rng.choice([12, 25, 38, 64, 10 ** 6]))
        Mlen, Nlen = (len(A), len(B)) if reverse else (len(A) - 1, len(B) - 1)
        sc, a, b, ops, cells = oracle.align_ex(A, B, x, reverse=reverse)
        want = oracle.semi_gapped(A, B, x, reverse=reverse, M=Mlen, N=Nlen)
        assert (sc, a, b, cells) == want, trial
        got, ca, cb = rescore(A, B, ops, reverse, Mlen, Nlen)
        assert (got, ca, cb) == (sc, a, b), trial


def test_no_xdrop_path_is_optimal():
    rng = np.random.default_rng(23)
    for trial in range(80):
        A, B = indel_pair(rng, int(rng.integers(2, 16)))
        sc, a, b, ops, _ = oracle.align_ex(A, B, 10 ** 6, reverse=True)
        assert (sc, a, b) == gotoh_anchored(A[::-1], B[::-1])
        assert rescore(A, B, ops, True, len(A), len(B))[0] == sc


def test_traceback_subject_self_and_indel():
    """Query = subject: one HSP over the whole sequence, every column an identity.  A homolog with
    an inserted block: the traced HSP's runs re-score to its score and count its gaps."""
    rng = np.random.default_rng(24)
    q = rand_seq(rng, 120)
    pre, _ = oracle.prelim_subject(q, q)
    tr, runs = oracle.traceback_subject(q, q, pre, cutoff=0)
    assert len(tr) == 1 and (tr[0]["q_off"], tr[0]["q_end"], tr[0]["num_ident"], tr[0]["gaps"]) == (0, 120, 120, 0)
    assert list(runs) == [(oracle.OP_SUB << 28) | 120]
    s = np.concatenate([rand_seq(rng, 30), mutate(rng, q[:60], 0.9), rand_seq(rng, 4), mutate(rng, q[60:], 0.9),
                        rand_seq(rng, 30)])
    pre, _ = oracle.prelim_subject(q, s)
    tr, runs = oracle.traceback_subject(q, s, pre, cutoff=0)
    h = tr[0]
    ops = np.concatenate([np.full(r & 0x0FFFFFFF, r >> 28, np.uint8)
                          for r in runs[h["op_first"]:h["op_first"] + h["op_count"]]])
    # re-score left to right from (q_off, s_off)
    qi, si, score, prev = int(h["q_off"]), int(h["s_off"]), 0, None
    for op in ops:
        if op == oracle.OP_SUB:
            score += int(M[q[qi], s[si]])
            qi += 1
            si += 1
        else:
            score -= 1 + (11 if op != prev else 0)
            if op == oracle.OP_GAP_Q:
                si += 1
            else:
                qi += 1
        prev = op
    assert (score, qi, si) == (int(h["score"]), int(h["q_end"]), int(h["s_end"]))
    assert h["gaps"] >= 4 and h["gap_opens"] >= 1 and h["align_len"] == len(ops)


def test_traceback_skips_contained_hsps():
    """A prelim HSP inside one already traced, with no higher score, is not aligned again (NCBI
    s_HSPIsContained in Blast_TracebackFromHSPList): the same result and DP cells without it."""
    rng = np.random.default_rng(25)
    q = rand_seq(rng, 100)
    pre, _ = oracle.prelim_subject(q, q)
    inner = pre.copy()
    inner["score"] -= 5
    inner["q_off"] += 10
    inner["s_off"] += 10
    inner["q_gs"] += 20
    inner["s_gs"] += 20
    both = np.concatenate([pre, inner])
    cells = []
    for lst in (pre, both):
        tr, runs = oracle.traceback_subject(q, q, lst, cutoff=0)
        cells.append((tr.tolist(), runs.tolist()))
    assert cells[0] == cells[1]


# NCBIstdaa codes of the letters the known-answer cases use
A_, D_, E_, G_, I_, M_, W_ = 1, 4, 5, 7, 9, 12, 20


def test_tie_rules_known_answers():
    """Equal-scoring alternatives decided by ALIGN_EX's tie rules, each worked by hand (BLOSUM62,
    gap cost 11 + 1 per letter, no X-drop).  All four cases run on reversed prefixes, so the rows
    are A read backwards (A') and the columns B read backwards (B'); ops are listed from the best
    cell back to the origin: 0 aligned pair, 1 B letter against a gap, 2 A letter against a gap."""
    def ops(A, B):
        sc, a, b, o, _ = oracle.align_ex(np.array(A, np.uint8), np.array(B, np.uint8), 10 ** 6, reverse=True)
        return sc, a, b, o.tolist()

    # 1. Diagonal against the column gap.  A' = A A M W (rows), B' = W M W.  Cell (2,1): the diagonal
    #    from (1,0) scores H(1,0) + s(A,W) = -12 - 3 = -15; the column gap opened below (1,1) scores
    #    H(1,1) - 12 = s(A,W) - 12 = -15.  Equal, and the diagonal is kept unless a gap is strictly
    #    greater, so the script from the origin is A/-, A/W, M/M, W/W (score -12 - 3 + 5 + 11 = 1).
    assert ops([W_, M_, A_, A_], [W_, M_, W_]) == (1, 4, 3, [0, 0, 0, 2])
    # 2. Diagonal against the row gap.  A' = M W M, B' = W A W M.  Cell (1,2): the diagonal from
    #    (0,1) scores -12 + s(M,A) = -13; the row gap opened after (1,1) scores s(M,W) - 12 = -13.
    #    Equal: the diagonal.  Script from the origin: -/W, M/A, W/W, M/M (-12 - 1 + 11 + 5 = 3).
    assert ops([M_, W_, M_], [M_, W_, A_, W_]) == (3, 3, 4, [0, 0, 0, 1])
    # 3. Row gap: extend or open.  A' = W I A W I, B' = A I W M W A W.  H(2,2) = s(W,A) + s(I,I) =
    #    1; the row gap after it enters (2,4) at 1 - 12 - 1 = -12, so extending it past (2,4) gives
    #    -13.  H(2,4) = H(1,3) + s(I,M) = (-13 + 11) + 1 = -1, so opening there gives -1 - 12 = -13.
    #    Equal, and a gap that can extend does (the extension bit is set when the extended score
    #    is at least the opened one): one 3-letter gap W M W after I/I, not two gaps.  Script from
    #    the origin: W/A, I/I, -/W, -/M, -/W, A/A, W/W (-3 + 4 - 14 + 4 + 11 = 2).
    assert ops([I_, W_, A_, I_, W_], [W_, A_, W_, M_, W_, I_, A_]) == (2, 4, 7, [0, 0, 1, 1, 1, 0, 0])
    # 4. Column gap: extend or open.  A' = I E W G W G W G D, B' = W W G D.  H(3,1) = -13 + 11 = -2
    #    (I, E against gaps, then W/W); its column gap enters (5,1) at -2 - 12 - 1 = -15 and would
    #    leave at -16.  H(5,1) = H(4,0) + s(W,W) = -15 + 11 = -4, opening at -4 - 12 = -16.  Equal:
    #    the gap extends, so G W G form one gap.  Script from the origin: I/-, E/-, W/W, G/-, W/-,
    #    G/-, W/W, G/G, D/D (-13 + 11 - 14 + 11 + 6 + 6 = 7).
    assert ops([D_, G_, W_, G_, W_, G_, W_, E_, I_], [D_, G_, W_, W_]) == (7, 9, 4, [0, 0, 0, 2, 2, 2, 0, 2, 2])


def test_clean_insertion_known_answer():
    """A subject that is the query with five W inserted after its 30th letter (neither neighbour a
    W): one HSP over both sequences whose script is 30 aligned pairs, the five subject letters
    against one gap, 30 aligned pairs; 60 identities, 65 columns, 5 gap letters, one opening."""
    rng = np.random.default_rng(26)
    while True:
        q = rand_seq(rng, 60)
        if W_ not in (q[29], q[30]):
            break
    s = np.concatenate([q[:30], np.full(5, W_, np.uint8), q[30:]])
    pre, _ = oracle.prelim_subject(q, s)
    tr, runs = oracle.traceback_subject(q, s, pre, cutoff=0)
    assert len(tr) == 1
    h = tr[0]
    assert (h["q_off"], h["q_end"], h["s_off"], h["s_end"]) == (0, 60, 0, 65)
    assert (h["num_ident"], h["align_len"], h["gaps"], h["gap_opens"]) == (60, 65, 5, 1)
    assert runs.tolist() == [(oracle.OP_SUB << 28) | 30, (oracle.OP_GAP_Q << 28) | 5, (oracle.OP_SUB << 28) | 30]


def test_containment_is_inclusive_and_skips_the_alignment():
    """NCBI s_HSPIsContained: an HSP whose score is at most a traced one's and whose end points lie
    inside its box (bounds included) is not aligned again -- the DP cell count does not change; one
    scoring a single point more is aligned."""
    rng = np.random.default_rng(27)
    q = rand_seq(rng, 90)
    s = np.concatenate([rand_seq(rng, 10), mutate(rng, q, 0.8), rand_seq(rng, 10)])
    pre, _ = oracle.prelim_subject(q, s)
    top = pre[:1]
    t, _, c1 = oracle.traceback_subject(q, s, top, return_cells=True)
    inner = top.copy()
    for f in ("score", "q_off", "q_end", "s_off", "s_end"):
        inner[f] = t[f][0]  # the traced box itself, bounds included, equal score
    inner["q_gs"], inner["s_gs"] = t["q_off"][0] + 3, t["s_off"][0] + 3
    t2, r2, c2 = oracle.traceback_subject(q, s, np.concatenate([top, inner]), return_cells=True)
    assert c2 == c1 and t2.tolist() == t.tolist()
    inner["score"] += 1
    _, _, c3 = oracle.traceback_subject(q, s, np.concatenate([top, inner]), return_cells=True)
    assert c3 > c1


def test_cutoff_is_inclusive():
    rng = np.random.default_rng(28)
    q = rand_seq(rng, 80)
    s = mutate(rng, q, 0.7)
    pre, _ = oracle.prelim_subject(q, s)
    t, _ = oracle.traceback_subject(q, s, pre[:1])
    sc = int(t["score"][0])
    assert len(oracle.traceback_subject(q, s, pre[:1], cutoff=sc)[0]) == 1
    assert len(oracle.traceback_subject(q, s, pre[:1], cutoff=sc + 1)[0]) == 0
