"""Pins the CPU oracle before it is trusted as the parity reference.

Each check is chosen so that a plausible slip in the oracle (a dropped term, an off-by-one in an
offset, a wrong stop rule, a transposed operand) fails: fixtures from the reference's own files,
closed forms, brute force on small inputs, and independent formulations of the same rule.
"""
import pathlib

import numpy as np
import pytest

import oracle

GOLDEN = pathlib.Path(__file__).parent / "golden"
M = oracle.matrix()


def fixture_matrix():
    return np.array([list(map(int, l.split())) for l in (GOLDEN / "blosum62_ncbistdaa.txt").read_text().split("\n") if l])


def fixture_queries():
    out = {}
    for line in (GOLDEN / "queries_ncbistdaa.txt").read_text().splitlines():
        name, *codes = line.split()
        out[name] = np.array(list(map(int, codes)), dtype=np.uint8)
    return out


def rand_seq(rng, n):
    # the 20 standard residues in NCBIstdaa
    std = np.array([1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22], np.uint8)
    return rng.choice(std, n).astype(np.uint8)


def mutate(rng, s, ident):
    t = s.copy()
    flip = rng.random(len(s)) > ident
    t[flip] = rand_seq(rng, int(flip.sum()))
    return t


# ---------------------------------------------------------------- matrix and neighbourhoods
def test_matrix_matches_reference_table():
    ref = fixture_matrix()  # BLAZE_gpu/util.hpp:30-57
    assert ref.shape == (28, 28)
    assert np.array_equal(M[1:], ref[1:])
    assert (M[0] == -32768).all() and (M[:, 0] == -32768).all()
    assert np.array_equal(M, M.T)
    assert np.array_equal(M[24], M[3])  # U scores as C


def brute_neighbors(w, T=11):
    x, y, z = w // 784, (w // 28) % 28, w % 28
    sc = M[x][:, None, None] + M[y][None, :, None] + M[z][None, None, :]
    u = np.nonzero(sc.reshape(-1) >= T)[0]
    return np.union1d(u, [w]).astype(np.int32)


def test_neighbors_brute_force():
    rng = np.random.default_rng(1)
    words = list(rng.integers(0, 28 ** 3, 300)) + [0, 28 ** 3 - 1, 1 * 784 + 1 * 28 + 1, 20 * 784 + 20 * 28 + 20]
    for w in words:
        assert np.array_equal(oracle.neighbors(int(w)), brute_neighbors(int(w))), w


ROBINSON = {1: 78.05, 3: 19.25, 4: 53.64, 5: 62.95, 6: 38.56, 7: 73.77, 8: 21.99, 9: 51.42, 10: 57.44,
            11: 90.19, 12: 22.43, 13: 44.87, 14: 52.03, 15: 42.64, 16: 51.29, 17: 71.20, 18: 58.41,
            19: 64.41, 20: 13.30, 22: 32.16}


def test_neighborhood_statistics():
    """Figures computed independently from the BLOSUM62 text: over the 27^3 non-gap words the
    largest |N(w)| (self included) is 308 and 8,589 words have |N| > 32; weighted by the
    Robinson-Robinson background the mean |N| is 33.8 and a random subject word hits a random
    query word with probability 0.00210."""
    sizes = np.array([len(oracle.neighbors(a * 784 + b * 28 + c)) for a in range(1, 28) for b in range(1, 28)
                      for c in range(1, 28)]).reshape(27, 27, 27)
    assert sizes.max() == 308
    assert (sizes > 32).sum() == 8589
    p = np.zeros(28)
    for k, v in ROBINSON.items():
        p[k] = v
    p /= p.sum()
    P = p[1:, None, None] * p[None, 1:, None] * p[None, None, 1:]
    assert round(float((sizes * P).sum()), 1) == 33.8
    hit = 0.0
    for a in ROBINSON:
        for b in ROBINSON:
            for c in ROBINSON:
                nb = oracle.neighbors(a * 784 + b * 28 + c)
                hit += p[a] * p[b] * p[c] * (p[nb // 784] * p[(nb // 28) % 28] * p[nb % 28]).sum()
    assert round(hit, 5) == 0.00210


def test_lookup_table_brute_force():
    rng = np.random.default_rng(2)
    q = rand_seq(rng, 60)
    lk = oracle.LookupTable(q)
    words = q[:-2].astype(int) * 784 + q[1:-1].astype(int) * 28 + q[2:].astype(int)
    for u in list(rng.integers(0, 28 ** 3, 400)) + list(words):
        want = [qi for qi, w in enumerate(words) if u in set(brute_neighbors(int(w)).tolist())]
        got = lk.entries[lk.prefix[u]:lk.prefix[u + 1]].tolist()
        assert got == want


# ---------------------------------------------------------------- ungapped X-drop
def test_extend_left_infinite_dropoff_is_best_prefix():
    """With no X-drop the left walk returns the best suffix sum ending at the start (or 0) and
    the shortest length reaching it."""
    rng = np.random.default_rng(3)
    for _ in range(300):
        n = int(rng.integers(1, 40))
        q, s = rand_seq(rng, n), rand_seq(rng, n)
        sc, length = oracle.extend_left(q, s, n - 1, n - 1, dropoff=10 ** 6)
        steps = M[q[::-1], s[::-1]].cumsum()
        best = max(0, steps.max())
        assert sc == best
        want_len = 0 if best == 0 else int(np.argmax(steps == best)) + 1
        assert length == want_len


def test_extend_right_stops_at_nonpositive_running_score():
    rng = np.random.default_rng(4)
    for _ in range(300):
        n = int(rng.integers(1, 40))
        q, s = rand_seq(rng, n), rand_seq(rng, n)
        start = int(rng.integers(0, 30))
        sc, length, last = oracle.extend_right(q, s, 0, 0, dropoff=10 ** 6, maxscore=start)
        run = start + M[q, s].cumsum()
        stop = np.nonzero(run <= 0)[0]
        end = int(stop[0]) if len(stop) else n
        seg = run[: end + 1] if len(stop) else run
        best = max(start, seg.max())
        assert sc == best
        assert length == (0 if best == start else int(np.argmax(seg == best)) + 1)
        assert last == (end if len(stop) else n)


def test_xdrop_monotone_in_dropoff():
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = 60
        q = rand_seq(rng, n)
        s = mutate(rng, q, 0.4)
        prev = -1
        for x in (4, 8, 16, 32, 10 ** 6):
            sc, _ = oracle.extend_left(q, s, n - 1, n - 1, dropoff=x)
            assert sc >= prev
            prev = sc


# ---------------------------------------------------------------- gapped start
def test_gapped_start_cases():
    rng = np.random.default_rng(6)
    for _ in range(300):
        n = int(rng.integers(1, 60))
        q, s = rand_seq(rng, n + 5), rand_seq(rng, n + 5)
        q0, s0 = 2, 3
        length = n
        got = oracle.gapped_start(q, s, q0, s0, length)
        if length <= 11:
            assert got == q0 + length // 2
            continue
        sc = M[q[q0:q0 + length], s[s0:s0 + length]]
        win = np.convolve(sc, np.ones(11, int), "valid")  # win[k] = sum sc[k..k+10]
        k = int(np.argmax(win == win.max()))
        if win.max() > 0:
            assert got == q0 + k + 10 - 5
        else:
            assert got == q0


# ---------------------------------------------------------------- semi-gapped DP
def gotoh_anchored(A, B, go=11, ge=1):
    """Full affine DP anchored at (0,0), free end: max over cells of the best alignment of
    A[:i] with B[:j] (gap of length k costs go + k*ge); returns (score, i, j) of the first
    row-major cell reaching the max."""
    m, n = len(A), len(B)
    NEG = -10 ** 9
    H = np.full((m + 1, n + 1), NEG, dtype=np.int64)
    E = np.full((m + 1, n + 1), NEG, dtype=np.int64)  # gap in B (consumes A)
    F = np.full((m + 1, n + 1), NEG, dtype=np.int64)
    H[0, 0] = 0
    for j in range(1, n + 1):
        F[0, j] = -(go + j * ge)
        H[0, j] = F[0, j]
    for i in range(1, m + 1):
        E[i, 0] = -(go + i * ge)
        H[i, 0] = E[i, 0]
        for j in range(1, n + 1):
            E[i, j] = max(E[i - 1, j] - ge, H[i - 1, j] - go - ge)
            F[i, j] = max(F[i, j - 1] - ge, H[i, j - 1] - go - ge)
            H[i, j] = max(H[i - 1, j - 1] + M[A[i - 1], B[j - 1]], E[i, j], F[i, j])
    best, bi, bj = 0, 0, 0
    for i in range(m + 1):
        for j in range(n + 1):
            if H[i, j] > best:
                best, bi, bj = int(H[i, j]), i, j
    return best, bi, bj


@pytest.mark.parametrize("reverse", [False, True])
def test_semi_gapped_no_xdrop_equals_gotoh(reverse):
    rng = np.random.default_rng(7)
    for _ in range(60):
        m, n = int(rng.integers(1, 18)), int(rng.integers(1, 18))
        A = rand_seq(rng, m)
        B = mutate(rng, A, 0.5)[:n] if n <= m else np.concatenate([mutate(rng, A, 0.5), rand_seq(rng, n - m)])
        if reverse:
            sc, a, b, _ = oracle.semi_gapped(A, B, 10 ** 7, reverse=True, M=m, N=len(B))
            want = gotoh_anchored(A[::-1], B[::-1])
        else:  # forward mode aligns A[1:], B[1:] (index 0 is the anchor letter)
            A1, B1 = np.concatenate([[1], A]).astype(np.uint8), np.concatenate([[1], B]).astype(np.uint8)
            sc, a, b, _ = oracle.semi_gapped(A1, B1, 10 ** 7, M=m, N=len(B))
            want = gotoh_anchored(A, B)
        assert (sc, a, b) == want


def clean_xdrop(A, B, x, go=11, ge=1, dead_gap_to_neg=False):
    """Independent formulation of the banded X-drop: each row computed as vectors with a prefix
    max for the horizontal gaps and for the running best, band bounds from the surviving cells.
    dead_gap_to_neg also drops the vertical-gap score of a cell that fails the X-drop (the
    16-bit LDS storage of the GPU kernel does this)."""
    goe = go + ge
    NEG = -(1 << 30)
    m, n = len(A), len(B)
    H = [0]
    Eg = [-goe]
    sc = -goe
    while len(H) <= n and sc >= -x:
        H.append(sc)
        Eg.append(sc - goe)
        sc -= ge
    first, size, best, boff = 0, len(H), 0, (0, 0)
    H += [NEG] * (n + 2 - len(H))
    Eg += [NEG] * (n + 2 - len(Eg))
    for a in range(1, m + 1):
        row = np.arange(first, size)
        D = np.array([NEG if b == first else H[b - 1] + M[A[a - 1], B[b - 1]] for b in row], dtype=np.int64)
        E = np.array([Eg[b] for b in row], dtype=np.int64)
        DE = np.maximum(D, E)
        # F(b) = max over b' < b of DE(b') - goe - (b - 1 - b') * ge
        F = np.full(len(row), NEG, dtype=np.int64)
        run = NEG
        for k in range(len(row)):
            F[k] = run
            run = max(run - ge, DE[k] - goe)
        Hrow = np.maximum(DE, F)
        before = np.maximum.accumulate(np.concatenate([[best], Hrow[:-1]]))
        alive = before - Hrow <= x
        if not alive.any():
            break
        k_first = int(np.argmax(alive))
        last = int(row[len(row) - 1 - int(np.argmax(alive[::-1]))])
        rmax = Hrow[alive].max()
        if rmax > best:
            best = int(rmax)
            boff = (a, int(row[int(np.argmax(np.where(alive, Hrow, NEG) == rmax))]))
        newH, newE = list(H), list(Eg)
        for k, b in enumerate(row):
            if alive[k]:
                newH[b] = int(Hrow[k])
                newE[b] = int(max(Hrow[k] - goe, E[k] - ge))
            elif b > row[k_first]:
                newH[b] = NEG
                if dead_gap_to_neg:
                    newE[b] = NEG
        gap_end = max(run, NEG)
        H, Eg = newH, newE
        first = int(row[k_first])
        if last < size - 1:
            size = last + 1
        else:
            gr = gap_end
            while gr >= best - x and size <= n:
                H[size], Eg[size] = gr, gr - goe
                gr -= ge
                size += 1
        if size <= n:
            H[size], Eg[size] = NEG, NEG
            size += 1
    return best, boff[0], boff[1]


def test_semi_gapped_matches_clean_formulation():
    """The oracle's NCBI-order loop equals the row-vector formulation (prefix max for the row
    gap and for the running best) that a lane-parallel kernel can use."""
    rng = np.random.default_rng(8)
    for trial in range(150):
        m = int(rng.integers(5, 70))
        A = rand_seq(rng, m)
        B = mutate(rng, A, float(rng.uniform(0.2, 0.9)))
        if trial % 3 == 0:  # an indel
            k = int(rng.integers(1, m))
            B = np.concatenate([B[:k], rand_seq(rng, int(rng.integers(1, 6))), B[k:]])
        x = int(rng.choice([12, 20, 38, 60]))
        sc, a, b, _ = oracle.semi_gapped(A, B, x, reverse=True, M=len(A), N=len(B))
        want = clean_xdrop(A[::-1], B[::-1], x)
        assert (sc, a, b) == want, trial


def test_semi_gapped_dead_gap_scores_do_not_matter():
    """Dropping the gap score of cells that failed the X-drop leaves score and end point unchanged:
    such a score is below best - X and stays there, so it can decide no later cell."""
    rng = np.random.default_rng(14)
    for trial in range(150):
        m = int(rng.integers(5, 80))
        A = rand_seq(rng, m)
        B = mutate(rng, A, float(rng.uniform(0.2, 0.9)))
        if trial % 2 == 0:
            k = int(rng.integers(1, m))
            B = np.concatenate([B[:k], rand_seq(rng, int(rng.integers(1, 9))), B[k:]])
        x = int(rng.choice([12, 20, 38, 60]))
        sc, a, b, _ = oracle.semi_gapped(A, B, x, reverse=True, M=len(A), N=len(B))
        assert (sc, a, b) == clean_xdrop(A[::-1], B[::-1], x, dead_gap_to_neg=True), trial


def test_semi_gapped_xdrop_bounds():
    rng = np.random.default_rng(9)
    for _ in range(60):
        m = int(rng.integers(5, 30))
        A = rand_seq(rng, m)
        B = mutate(rng, A, 0.5)
        full, _, _ = gotoh_anchored(A[::-1], B[::-1])
        prev = -1
        for x in (12, 20, 38, 80, 10 ** 6):
            sc, _, _, _ = oracle.semi_gapped(A, B, x, reverse=True, M=m, N=m)
            assert sc <= full
            assert sc >= prev
            prev = sc
        assert prev == full


# ---------------------------------------------------------------- word finder
def diagonal_major_word_finder(q, s, p):
    """Same two-hit rule, visited diagonal by diagonal instead of subject position by position."""
    words_q = q[:-2].astype(int) * 784 + q[1:-1].astype(int) * 28 + q[2:].astype(int)
    lk = oracle.LookupTable(q)
    hits = {}
    for si in range(len(s) - 2):
        w = int(s[si]) * 784 + int(s[si + 1]) * 28 + int(s[si + 2])
        for qi in lk.entries[lk.prefix[w]:lk.prefix[w + 1]]:
            hits.setdefault(int(qi) - si, []).append((si, int(qi)))
    out = []
    for d, hs in hits.items():
        last, flag = -(10 ** 9), False
        for si, qi in sorted(hs):
            if flag:
                if si < last:
                    continue
                last, flag = si, False
                continue
            diff = si - last
            if diff >= p.window:
                last = si
                continue
            if diff < 3:
                continue
            # two-hit extension via the oracle's walks
            left_score, score, right_d = 0, 0, 0
            for i in range(3):
                score += M[q[qi + i], s[si + i]]
                if score > left_score:
                    left_score, right_d = score, i + 1
            qr, sr = qi + right_d, si + right_d
            ls, ld = oracle.extend_left(q, s, qr - 1, sr - 1, p.x_ungapped, 0)
            rs, rd, slast = 0, 0, sr
            rext = ld >= sr - (last + 3)
            if rext:
                rs, rd, slast = oracle.extend_right(q, s, qr, sr, p.x_ungapped, ls)
            sc = max(ls, rs)
            if sc >= p.ungapped_cutoff:
                out.append((qr - ld, sr - ld, ld + rd, sc))
            if rext:
                last, flag = slast - 2, True
            else:
                last = si
    return sorted(out)


def test_word_finder_order_independent_of_scan_direction():
    rng = np.random.default_rng(10)
    p = oracle.params()
    for _ in range(25):
        q = rand_seq(rng, int(rng.integers(20, 120)))
        s = np.concatenate([rand_seq(rng, 30), mutate(rng, q, 0.6), rand_seq(rng, 30)])
        got, _ = oracle.word_finder(q, s, p)
        got = sorted((int(h["q_start"]), int(h["s_start"]), int(h["len"]), int(h["score"])) for h in got)
        assert got == diagonal_major_word_finder(q, s, p)


def test_init_hit_scores_are_segment_sums():
    rng = np.random.default_rng(11)
    p = oracle.params()
    n = 0
    for _ in range(30):
        q = rand_seq(rng, 150)
        s = np.concatenate([rand_seq(rng, 40), mutate(rng, q, 0.55), rand_seq(rng, 40)])
        hits, _ = oracle.word_finder(q, s, p)
        for h in hits:
            seg = M[q[h["q_start"]:h["q_start"] + h["len"]], s[h["s_start"]:h["s_start"] + h["len"]]]
            assert seg.sum() == h["score"] >= p.ungapped_cutoff
            assert h["q_start"] - h["s_start"] == h["q_hit"] - h["s_hit"]  # on the hit's diagonal
            n += 1
    assert n > 20


# ---------------------------------------------------------------- whole prelim stage
def test_self_hit_closed_form():
    """A subject equal to the query aligns end to end on the main diagonal: one HSP over the
    whole sequence scoring the sum of the self scores."""
    for name, q in list(fixture_queries().items())[:12]:
        q = q[(q >= 1) & (q < 28)]
        hsps, _ = oracle.prelim_subject(q, q)
        top = hsps[0]
        assert top["score"] == int(M[q, q].sum()), name
        assert (top["q_off"], top["q_end"], top["s_off"], top["s_end"]) == (0, len(q), 0, len(q))


def test_hsp_list_invariants_and_recall():
    rng = np.random.default_rng(12)
    p = oracle.params()
    q = rand_seq(rng, 300)
    subs, planted = [], []
    for i in range(200):
        if i % 10 == 0:
            subs.append(np.concatenate([rand_seq(rng, 50), mutate(rng, q, 0.5), rand_seq(rng, 50)]))
            planted.append(i)
        else:
            subs.append(rand_seq(rng, int(rng.integers(30, 800))))
    offs = np.zeros(len(subs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    counts, hsps, st = oracle.prelim_db(q, np.concatenate(subs), offs, p, nthreads=4)
    lists = oracle.as_lists(counts, hsps)
    assert sum(1 for i in planted if i in lists) == len(planted)
    for oid, h in lists.items():
        sc = h["score"]
        assert (np.diff(sc) <= 0).all() and (sc >= p.gapped_cutoff).all()
        starts = set(zip(h["q_off"].tolist(), h["s_off"].tolist()))
        ends = set(zip(h["q_end"].tolist(), h["s_end"].tolist()))
        assert len(starts) == len(h) and len(ends) == len(h)
        assert (h["q_off"] <= h["q_gs"] + 1).all() and (h["q_end"] >= h["q_gs"] + 1).all()
        assert (h["q_end"] <= len(q)).all() and (h["s_end"] <= len(subs[oid])).all()
    assert st["hsps"] == len(hsps)


def test_threads_do_not_change_results():
    rng = np.random.default_rng(13)
    q = rand_seq(rng, 250)
    subs = [np.concatenate([rand_seq(rng, 20), mutate(rng, q, 0.45)]) if i % 7 == 0 else rand_seq(rng, 300)
            for i in range(300)]
    offs = np.zeros(len(subs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    res = np.concatenate(subs)
    a = oracle.prelim_db(q, res, offs, nthreads=1)
    b = oracle.prelim_db(q, res, offs, nthreads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
