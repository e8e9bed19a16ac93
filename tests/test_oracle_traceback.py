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
