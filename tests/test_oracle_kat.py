"""Known-answer tests that pin the oracle's boundary rules one at a time.

Each case is built so that exactly one rule decides the outcome, and the expected answer is
derived by hand in the docstring from the rule as the reference states it (BLAZE_gpu/gpuBLAZE.cu)
or, where the rule lives only in NCBI BLAST+ 2.13.0 (not vendored by the reference), from NCBI's
published routine named in the docstring.  Letter codes are NCBIstdaa: A 1, C 3, E 5, W 20;
BLOSUM62 (BLAZE_gpu/util.hpp:30-57): W/W 11, E/E 5, C/E -4, and the gap letter 0 scores -32768
against everything, so an ungapped walk stops on it and no word containing it is ever a hit.
"""
import numpy as np

import oracle

M = oracle.matrix()
A, C, E, W = 1, 3, 5, 20
STD = np.array([1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22], np.uint8)


def std_seq(seed, n):
    """n standard residues (self scores 4..11, all positive)."""
    return np.random.default_rng(seed).choice(STD, n).astype(np.uint8)


def self_score(x):
    return int(M[x, x].sum())


def diag0_hits(q, s, p):
    """Init hits of the word finder on diagonal q - s = 0, as (q_hit, q_start, len, score)."""
    hits, _ = oracle.word_finder(q, s, p)
    return sorted((int(h["q_hit"]), int(h["q_start"]), int(h["len"]), int(h["score"]))
                  for h in hits if h["q_hit"] == h["s_hit"])


def word_positions(s):
    """Subject offsets whose word has no gap letter (the only words that can hit)."""
    return [i for i in range(len(s) - 2) if s[i] and s[i + 1] and s[i + 2]]


# ------------------------------------------------------------------ ungapped X-drop walks
def test_left_walk_stops_when_drop_equals_xdrop():
    """gpuBLAZE.cu:727 stops the left walk once (best - running) >= dropoff.
    q = W C E, s = W E E, walking left from offset 2: E/E +5 (best 5, 1 letter), C/E -4 (running
    1, drop 4), W/W +11 (running 12).  X = 4: the drop reaches 4 at the second letter -> (5, 1).
    X = 5: the walk goes on and the third letter makes a new best -> (12, 3)."""
    q = np.array([W, C, E], np.uint8)
    s = np.array([W, E, E], np.uint8)
    assert oracle.extend_left(q, s, 2, 2, dropoff=4) == (5, 1)
    assert oracle.extend_left(q, s, 2, 2, dropoff=5) == (12, 3)


def test_right_walk_stops_on_drop_and_on_nonpositive_score():
    """gpuBLAZE.cu:826 stops the right walk once running <= 0 or (best - running) >= dropoff.
    Start score 10, q = E C W, s = E E W: +5 (15, best, 1 letter), -4 (11, drop 4), +11 (22).
    X = 4 -> stops at letter 1: (15, 1, last 1); X = 5 -> (22, 3, last 3).
    Start score 4, q = C E W, s = E E W: -4 makes the running score 0 -> stop at once: (4, 0, 0)."""
    q = np.array([E, C, W], np.uint8)
    s = np.array([E, E, W], np.uint8)
    assert oracle.extend_right(q, s, 0, 0, dropoff=4, maxscore=10) == (15, 1, 1)
    assert oracle.extend_right(q, s, 0, 0, dropoff=5, maxscore=10) == (22, 3, 3)
    q2 = np.array([C, E, W], np.uint8)
    assert oracle.extend_right(q2, s, 0, 0, dropoff=16, maxscore=4) == (4, 0, 0)


# ------------------------------------------------------------------ two-hit word finder
def _isolated_words(q, starts, n=3):
    """Subject of q's length that is all gap letters except q's letters at [a, a + n) for a in starts."""
    s = np.zeros(len(q), np.uint8)
    for a in starts:
        s[a:a + n] = q[a:a + n]
    return s


def test_two_hit_window_is_exclusive():
    """gpuBLAZE.cu:1033: a hit window or more letters past the diagonal's last hit starts over
    (it becomes the last hit); a closer one (and >= 3 past it) is extended.  With window 40 and
    the only two words of the subject at offsets 10 and 10 + d on diagonal 0 (each an exact
    query word, everything else gap letters): d = 39 -> one extension at the second word; d = 40
    -> none.  The extension's left walk covers the word (3 letters, self scores > 0) and stops on
    the gap letter before it, short of the first hit, so there is no right walk: the init hit is
    the word itself, scoring its self score (cutoff 0 keeps every extension)."""
    q = std_seq(1, 100)
    p = oracle.params(ungapped_cutoff=0)
    for d, want in ((39, True), (40, False)):
        s = _isolated_words(q, [10, 10 + d])
        assert word_positions(s) == [10, 10 + d]
        b = 10 + d
        expect = [(b, b, 3, self_score(q[b:b + 3]))] if want else []
        assert diag0_hits(q, s, p) == expect, d


def test_two_hit_overlap_needs_three_letters():
    """gpuBLAZE.cu:1039: a hit fewer than 3 letters (the word size) past the last hit is dropped.
    Six copied letters at offset 20 give diagonal-0 hits at 20 (last hit), 21, 22 (dropped: 1 and
    2 past) and 23 (3 past: extended).  Its left walk covers the six letters and stops on the gap
    letter at 19, reaching the first hit, so the right walk runs and stops at once on the gap
    letter at 26: the init hit is the six letters, scoring their self scores.  Five copied
    letters give hits at 20, 21, 22 only: no extension."""
    q = std_seq(2, 80)
    p = oracle.params(ungapped_cutoff=0)
    s6 = _isolated_words(q, [20], n=6)
    assert word_positions(s6) == [20, 21, 22, 23]
    assert diag0_hits(q, s6, p) == [(23, 20, 6, self_score(q[20:26]))]
    s5 = _isolated_words(q, [20], n=5)
    assert word_positions(s5) == [20, 21, 22]
    assert diag0_hits(q, s5, p) == []


def test_two_hit_flag_after_right_extension():
    """NCBI's two-hit rule (aa_ungapped.c s_BlastAaWordFinder_TwoHit, called through
    blast_engine.c:731): after an extension that ran its right walk, the diagonal's last hit is
    set to (where the right walk stopped) - 2 with a flag; while the flag is set, hits before that
    are dropped and the first at or past it becomes the last hit (no extension) and clears it.
    q: 40 standard letters; s = q with a gap letter at 20.  Diagonal 0 has hits at every offset
    except 18, 19, 20.  0: last hit.  1, 2: too close.  3: extended; the left walk reaches offset
    0, the right walk stops on the gap letter (stop 20): init hit [0, 20), last = 18, flag.
    4..17: before 18, dropped.  21: first hit past 18 -> last hit, flag cleared, no extension.
    22, 23: too close.  24: extended; left walk back to 21 (gap letter at 20), right walk to the
    end (stop 40): init hit [21, 40), last = 38, flag; 25..37 dropped."""
    q = std_seq(3, 40)
    s = q.copy()
    s[20] = 0
    p = oracle.params(ungapped_cutoff=0)
    assert diag0_hits(q, s, p) == [(3, 0, 20, self_score(q[0:20])), (24, 21, 19, self_score(q[21:40]))]


def test_two_hit_flag_barrier_and_exact_reach_with_ambiguity_letters():
    """Both rules need a hit that an X-drop walk stops at, which only words of ambiguity letters
    allow: X scores -1 against itself, yet XXX hits as the query's own word (the exact word is
    always in its neighbourhood, prelim_stage.cpp:380).  The right walk runs when the left walk
    reached the last hit's end, left length >= (second hit's right start) - (last hit + 3)
    (gpuBLAZE.cu:1143); the first hit at or past a flagged barrier becomes the last hit.
    q = s = 10 standard letters, 20 X, 10 standard letters; diagonal 0 hits everywhere.
    0: last.  3: extended; left walk to 0, right walk over the standard letters (best at 9) into
    the X run, where the drop reaches 16 at letter 25: init hit [0, 10), last = 23, flag.
    4..22 dropped.  23 (XXX, exactly the barrier): last hit, flag cleared.  24, 25: too close.
    26 (XXX): extended; its word has no positive prefix (right start 26), the left walk never
    gains (left length 0) and 0 >= 26 - (23 + 3): the right walk runs and stops at once (running
    score 0): init hit of length 0 and score 0, last = 24, flag.  27: last hit, flag cleared.
    28, 29: too close.  30: extended; left walk 30..32 (the X run only loses), 3 >= 33 - 30, right
    walk to the end: init hit [30, 40)."""
    X = 21
    q = np.concatenate([std_seq(6, 10), np.full(20, X, np.uint8), std_seq(7, 10)]).astype(np.uint8)
    p = oracle.params(ungapped_cutoff=0)
    assert diag0_hits(q, q.copy(), p) == [(3, 0, 10, self_score(q[0:10])), (26, 26, 0, 0),
                                           (30, 30, 10, self_score(q[30:40]))]


def test_ungapped_cutoff_is_inclusive():
    """An extension scoring exactly the ungapped cutoff is kept (score >= cutoff, NCBI
    BlastSaveInitHsp; the reference passes the cutoff at gpu_cpu_common.h:43).  The window case's
    init hit scores the self score S of its word: cutoff S keeps it, S + 1 drops it."""
    q = std_seq(1, 100)
    s = _isolated_words(q, [10, 49])
    S = self_score(q[49:52])
    assert diag0_hits(q, s, oracle.params(ungapped_cutoff=S)) == [(49, 49, 3, S)]
    assert diag0_hits(q, s, oracle.params(ungapped_cutoff=S + 1)) == []


# ------------------------------------------------------------------ gapped stage
def _self_case(n=60, seed=4):
    q = std_seq(seed, n)
    return q, q.copy(), self_score(q)


def test_self_alignment_box():
    """Query = subject: the gapped extension from any init hit on the main diagonal is the whole
    diagonal (every self score is positive and every off-diagonal pair scores less), box
    [0, n) x [0, n), score = sum of self scores; a short init hit's gapped start is its middle."""
    q, s, S = _self_case()
    h, st = oracle.gapped_stage(q, s, [(2, 2, 5, 30)])
    assert [tuple(x) for x in h] == [(S, 0, 60, 0, 60, 4, 4)]
    assert st["gapped_ext"] == 1


def test_containment_equal_score_is_skipped():
    """NCBI s_HSPIsContained (blast_itree.c, through BlastIntervalTreeContainsHSP in
    BLAST_GetGappedScore): an init hit is skipped when its score is <= a saved HSP's score and
    both its end points lie in that HSP's box.  Init hits A (2, 2, len 5) and B (20, 20, len 5),
    both scoring S (A first: equal scores, lower subject start); A's HSP is the whole diagonal
    with score S, so B (score S <= S, inside) is skipped: one gapped extension.  With B scoring
    S + 1 it would come first instead; the order rule is pinned below."""
    q, s, S = _self_case()
    h, st = oracle.gapped_stage(q, s, [(2, 2, 5, S), (20, 20, 5, S)])
    assert st["gapped_ext"] == 1
    assert [tuple(x) for x in h] == [(S, 0, 60, 0, 60, 4, 4)]


def test_containment_uses_inclusive_segment_end():
    """The init hit's segment enters the containment test with its last letter as its end
    (q_start + len - 1, the convention of NCBI BLAST_GetGappedScore's tmp_hsp, where a point hit
    has start == end).  HSP box [0, 40] from A; B = (36, 36, len 5) ends on letter 40 = the box's
    end: contained, skipped.  (Had the end been taken as q_start + len = 41, B would be
    extended.)  The subject is the query's first 40 letters and a gap letter, its last: A's HSP
    can gain nothing past letter 39, so its box is [0, 40] x [0, 40]."""
    q = std_seq(5, 60)
    s = np.concatenate([q[:40], [0]]).astype(np.uint8)
    S = self_score(q[:40])
    h, st = oracle.gapped_stage(q, s, [(2, 2, 5, S), (36, 36, 5, S - 5)])
    assert [tuple(x)[:5] for x in h] == [(S, 0, 40, 0, 40)]
    assert st["gapped_ext"] == 1


def test_init_hit_order_subject_start_then_length_then_query_start():
    """NCBI score_compare_match (BLAST_GetGappedScore's sort): score desc, subject start asc,
    length desc, query start asc.  Query = subject (60 letters), so the first init hit taken
    yields the whole-diagonal HSP and contains the others; its gapped start shows which came
    first.
    - (30, 30, 5) and (5, 5, 5), equal scores: subject start 5 first -> gapped start 5 + 2 = 7.
    - (10, 10, 3) and (10, 10, 9): length 9 first -> gapped start 10 + 4 = 14.
    - (12, 10, 5) on diagonal 2 and (10, 10, 5) on diagonal 0, same subject start and length:
      query start 10 first -> the diagonal HSP with gapped start 12, and the diagonal-2 hit
      (box q 12..16, s 10..14, inside) is skipped: one extension."""
    q, s, S = _self_case()
    h, _ = oracle.gapped_stage(q, s, [(30, 30, 5, 10), (5, 5, 5, 10)])
    assert [tuple(x) for x in h] == [(S, 0, 60, 0, 60, 7, 7)]
    h, _ = oracle.gapped_stage(q, s, [(10, 10, 3, 10), (10, 10, 9, 10)])
    assert [tuple(x) for x in h] == [(S, 0, 60, 0, 60, 14, 14)]
    h, st = oracle.gapped_stage(q, s, [(12, 10, 5, 10), (10, 10, 5, 10)])
    assert [tuple(x) for x in h] == [(S, 0, 60, 0, 60, 12, 12)]
    assert st["gapped_ext"] == 1


def test_gapped_cutoff_is_inclusive():
    """An HSP scoring exactly the gapped cutoff is saved (score >= cutoff, NCBI
    BLAST_GetGappedScore; the reference passes the cutoff at gpu_cpu_common.h:43)."""
    q, s, S = _self_case()
    assert len(oracle.gapped_stage(q, s, [(2, 2, 5, 30)], oracle.params(gapped_cutoff=S))[0]) == 1
    assert len(oracle.gapped_stage(q, s, [(2, 2, 5, 30)], oracle.params(gapped_cutoff=S + 1))[0]) == 0


# ------------------------------------------------------------------ purge and final order
def rec(*rows):
    return oracle.hsp_records(rows)


def as_tuples(h):
    return [tuple(int(v) for v in x) for x in h]


def test_purge_common_start_keeps_best_by_score_then_longest_end():
    """NCBI Blast_HSPListPurgeHSPsWithCommonEndpoints (purge = TRUE for blastp) sorts by
    s_QueryOffsetCompareHSPs -- query start, subject start, then score desc, query end desc,
    subject end desc -- and keeps the first HSP of each group sharing (query start, subject
    start).  Records are (score, q_off, q_end, s_off, s_end, q_gs, s_gs)."""
    p1 = (50, 0, 30, 0, 30, 5, 5)
    assert as_tuples(oracle.purge(rec((50, 0, 25, 0, 31, 6, 6), p1))) == [p1]      # query end 30 > 25
    assert as_tuples(oracle.purge(rec((50, 0, 30, 0, 28, 7, 7), p1))) == [p1]      # subject end 30 > 28
    assert as_tuples(oracle.purge(rec(p1, (60, 0, 10, 0, 12, 1, 1)))) == [(60, 0, 10, 0, 12, 1, 1)]  # score


def test_purge_common_end_keeps_best_by_score_then_shortest_range():
    """The second pass sorts by s_QueryEndCompareHSPs -- query end, subject end, then score desc,
    query start desc, subject start desc (the shortest range first) -- and keeps the first HSP
    of each group sharing (query end, subject end)."""
    q1 = (50, 5, 40, 5, 40, 20, 20)
    assert as_tuples(oracle.purge(rec(q1, (50, 8, 40, 6, 40, 21, 21)))) == [(50, 8, 40, 6, 40, 21, 21)]
    assert as_tuples(oracle.purge(rec(q1, (50, 5, 40, 9, 40, 22, 22)))) == [(50, 5, 40, 9, 40, 22, 22)]
    assert as_tuples(oracle.purge(rec((40, 1, 40, 1, 40, 9, 9), q1))) == [q1]  # score first


def test_final_order_ties():
    """NCBI ScoreCompareHSPs (Blast_HSPListSortByScore, blast_engine.c:801): score desc, subject
    start asc, subject end desc, query start asc, query end desc.  Each pair differs in one key
    only and comes in the wrong order."""
    pairs = [
        ((50, 0, 20, 5, 30, 1, 1), (60, 0, 20, 5, 30, 1, 1)),    # score
        ((50, 0, 20, 6, 30, 1, 1), (50, 0, 20, 5, 30, 1, 1)),    # subject start asc
        ((50, 0, 20, 5, 30, 1, 1), (50, 0, 20, 5, 31, 1, 1)),    # subject end desc
        ((50, 2, 20, 5, 30, 1, 1), (50, 1, 20, 5, 30, 1, 1)),    # query start asc
        ((50, 0, 20, 5, 30, 1, 1), (50, 0, 25, 5, 30, 1, 1)),    # query end desc
    ]
    for wrong_first, right_first in pairs:
        assert as_tuples(oracle.sort_hsps(rec(wrong_first, right_first))) == [right_first, wrong_first]
