"""CPU restatement of the prelim stage's statistics -- test infrastructure only.

Only tests/ import this module, as the checker of the library's statistics (blz_stats.cpp).
The reference takes these values from NCBI BLAST+ 2.13.0 (blast_stat.c, blast_parameters.c,
blast_hits.c; not vendored): cutoffs at BLAZE_ncbi/blast_engine.c:2971-2981, e-values at :1441,
the prelim reap at :896-930,1457, the hit list at BLAZE_ncbi/prelim_stage.cpp:1462-1498.  Each
quantity is computed here by a different route from the library's where one exists:
  - lambda as the real root above 1 of the score polynomial sum_s P(s) x^s = 1 (x = e^lambda);
  - K by the Karlin-Altschul (1990) lattice formula with numpy convolutions;
  - the length adjustment by a linear scan for the largest integer satisfying its inequality;
  - cutoff and reap scores by scanning integer scores for the e-value condition.
"""
from __future__ import annotations

import math

import numpy as np

import oracle

ROBINSON = {1: 78.05, 3: 19.25, 4: 53.64, 5: 62.95, 6: 38.56, 7: 73.77, 8: 21.99, 9: 51.42, 10: 57.44,
            11: 90.19, 12: 22.43, 13: 44.87, 14: 52.03, 15: 42.64, 16: 51.29, 17: 71.20, 18: 58.41,
            19: 64.41, 20: 13.30, 22: 32.16}

# NCBI's published BLOSUM62 gapped statistics (lambda, K, H, alpha, beta) for gap costs 11/1
GAPPED_11_1 = (0.267, 0.041, 0.14, 1.9, -30.0)


def background() -> np.ndarray:
    p = np.zeros(28)
    for k, v in ROBINSON.items():
        p[k] = v
    return p / p.sum()


def score_distribution(freq=None):
    """P(s) over s = lo..hi for letters drawn from freq (default the background) against the
    background, and lo."""
    M = oracle.matrix()
    q = background() if freq is None else np.asarray(freq, float) / np.sum(freq)
    bg = background()
    lo, hi = -4, 11
    P = np.zeros(hi - lo + 1)
    for i in range(1, 28):
        for j in range(1, 28):
            w = q[i] * bg[j]
            if w > 0:
                P[M[i, j] - lo] += w
    return P, lo


def karlin(P, lo):
    """(lambda, K, H) of the score distribution P over lo.. (Karlin & Altschul 1990)."""
    # x = e^lambda solves sum_s P(s) x^(s - lo) - x^(-lo) = 0, a polynomial in x
    coef = P.copy()
    coef[-lo] -= 1.0
    roots = np.roots(coef[::-1])
    real = [r.real for r in roots if abs(r.imag) < 1e-9 and r.real > 1 + 1e-9]
    lam = math.log(min(real))
    s = np.arange(lo, lo + len(P))
    H = lam * float((P * s * np.exp(lam * s)).sum())
    d = int(np.gcd.reduce(np.abs(s[P > 0])))
    sigma, Pk, klo = 0.0, P.copy(), lo
    for k in range(1, 5000):
        sk = np.arange(klo, klo + len(Pk))
        term = float((Pk[sk < 0] * np.exp(lam * sk[sk < 0])).sum() + Pk[sk >= 0].sum())
        sigma += term / k
        if term / k < 1e-16 * sigma:
            break
        Pk = np.convolve(Pk, P)
        klo += lo
    K = lam * d * math.exp(-2 * sigma) / (H * (1 - math.exp(-lam * d)))
    return lam, K, H


def length_adjustment(K, lam, alpha, beta, m, n, N):
    """Largest integer l >= 0 with alpha/lam * (ln K + ln((m - l)(n - N l))) + beta >= l, scanning
    l upward while (m - l)(n - N l) stays above max(m, n) / K (0 if it never is)."""
    best = 0
    l = 0
    while True:
        ss = (m - l) * (n - N * l)
        if ss <= max(m, n) / K:
            break
        if alpha / lam * (math.log(K) + math.log(ss)) + beta >= l:
            best = l
        l += 1
    return best


def evalue(score, lam, K, searchsp):
    return searchsp * math.exp(-lam * score + math.log(K))


def setup(qlen, db_residues, db_seqs, expect=10.0, prelim_evalue=None):
    """The statistics of one query with BLOSUM62 11/1 and the default bit values."""
    lu, Ku, Hu = karlin(*score_distribution())
    lg, Kg, Hg, alpha, beta = GAPPED_11_1
    adj = length_adjustment(Kg, lg, alpha, beta, qlen, db_residues, db_seqs)
    ss = float(max(1, db_residues - db_seqs * adj)) * float(max(1, qlen - adj))
    prelim = prelim_evalue or expect
    # smallest integer scores whose e-value is within the thresholds
    gapped_cutoff = next(S for S in range(1, 10 ** 6) if evalue(S, lg, Kg, ss) <= max(expect, 1e-7))
    reap = next(S for S in range(1, 10 ** 6) if evalue(S, lg, Kg, ss) <= prelim)
    return dict(lambda_u=lu, K_u=Ku, H_u=Hu, lambda_g=lg, K_g=Kg, adj=adj, searchsp=ss,
                gapped_cutoff=gapped_cutoff, reap_score=reap,
                x_ungapped=math.ceil(7 * math.log(2) / lu), x_gapped=int(15 * math.log(2) / lg),
                x_final=int(25 * math.log(2) / lg), gap_trigger=int((22 * math.log(2) + math.log(Ku)) / lu))


def reap(lists: dict, lam, K, searchsp, prelim_evalue) -> dict:
    """Subject lists with the HSPs above the prelim e-value removed (empty lists dropped)."""
    out = {}
    for oid, h in lists.items():
        keep = h[np.array([evalue(int(s), lam, K, searchsp) <= prelim_evalue for s in h["score"]], bool)]
        if len(keep):
            out[oid] = keep
    return out


def hitlist(lists: dict, lam, K, searchsp, size):
    """(oid, best e-value) of the best `size` subjects in NCBI's list order: best e-value ascending
    (below 1e-180 counted equal), best score descending, higher ordinal id first."""
    rows = []
    for oid, h in lists.items():
        best = int(h["score"][0])
        ev = evalue(best, lam, K, searchsp)
        rows.append((0.0 if ev < 1e-180 else ev, -best, -oid, oid, ev))
    rows.sort()
    return [(r[3], r[4]) for r in rows[:size]]
