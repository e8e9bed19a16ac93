"""Search statistics of the library (blz_stats.cpp), on the CPU: no device is needed.

Pinned by NCBI's published BLOSUM62 figures (ungapped lambda 0.3176, K 0.134, H 0.4012 from the
Robinson-Robinson background), by the constants the reference hard-codes from the default bit
values (ungapped X-drop 16, gapped X-drop 38, gap trigger 41: BLAZE_gpu/gpuBLAZE.cu:37,59-63),
by the equation lambda solves, and by brute force for the length adjustment and the cutoffs.
The oracle's restatement (oracle/stats.py) computes lambda as a polynomial root and the
adjustment by a linear scan, independently of the library's bisection and bracketing."""
import math

import numpy as np
import pytest

import blz
import oracle
import stats as ostats


def test_ungapped_statistics_match_published_blosum62():
    k = blz.karlin_ungapped()
    assert round(k.lambda_, 4) == 0.3176 and round(k.K, 3) == 0.134 and round(k.H, 4) == 0.4012
    lam, K, H = ostats.karlin(*ostats.score_distribution())
    assert abs(k.lambda_ - lam) < 1e-10 and abs(k.K - K) < 1e-9 and abs(k.H - H) < 1e-9
    assert abs(k.logK - math.log(k.K)) < 1e-12


def test_lambda_solves_its_equation():
    """sum_ij p_i p_j exp(lambda s_ij) = 1 over the 20 standard residues."""
    k = blz.karlin_ungapped()
    M = oracle.matrix()
    p = ostats.background()
    total = sum(p[i] * p[j] * math.exp(k.lambda_ * M[i, j]) for i in range(1, 28) for j in range(1, 28)
                if p[i] > 0 and p[j] > 0)
    assert abs(total - 1.0) < 1e-12


def test_default_bits_reproduce_the_reference_constants():
    """7 bits -> ungapped X-drop 16 (gpuBLAZE.cu:59-60), 15 bits -> gapped X-drop 38 (:37),
    22 bits -> gap trigger 41 (the ungapped cutoff, :63), 25 bits -> traceback X-drop 64."""
    p, s = blz.search_setup(np.ones(500, np.uint8), 825_563_801, 2_000_000)
    assert (p.x_ungapped, p.x_gapped, s.gap_trigger, s.x_final) == (16, 38, 41, 64)
    assert p.ungapped_cutoff == min(41, p.gapped_cutoff)
    assert (p.word_threshold, p.window, p.gap_open, p.gap_extend) == (11, 40, 11, 1)
    o = ostats.setup(500, 825_563_801, 2_000_000)
    assert (o["x_ungapped"], o["x_gapped"], o["gap_trigger"], o["x_final"]) == (16, 38, 41, 64)


def test_gapped_table():
    k, a, b = blz.karlin_gapped(11, 1)
    assert (k.lambda_, k.K, k.H, a, b) == ostats.GAPPED_11_1
    with pytest.raises(blz.BlzError):
        blz.karlin_gapped(5, 5)


@pytest.mark.parametrize("m,n,N", [(300, 4_130_003, 10_000), (500, 825_563_801, 2_000_000), (50, 10 ** 6, 3000),
                                   (3000, 825_563_801, 2_000_000), (1500, 551_901_670, 500_000), (30, 5000, 10),
                                   (350, 16_512_781_762, 40_000_000), (4000, 16_512_781_762, 40_000_000),
                                   (10, 100, 1), (7095, 4_130_003, 10_000)])
def test_length_adjustment_and_cutoffs(m, n, N):
    """The length adjustment is the largest integer satisfying NCBI's fixed-point inequality (brute
    force); the gapped cutoff and the reap score are the lowest scores whose e-value is within E
    and the prelim threshold."""
    opts = blz.stats_options(prelim_evalue=1e-3)
    p, s = blz.search_setup(np.ones(m, np.uint8), n, N, opts)
    o = ostats.setup(m, n, N, expect=10.0, prelim_evalue=1e-3)
    assert s.length_adjustment == o["adj"]
    assert s.eff_searchsp == o["searchsp"]
    assert p.gapped_cutoff == o["gapped_cutoff"]
    assert p.reap_score == s.reap_score == o["reap_score"]
    ev = lambda S: blz.evalue(s, S)  # noqa: E731
    assert ev(p.gapped_cutoff) <= 10.0 < ev(p.gapped_cutoff - 1)
    assert ev(s.reap_score) <= 1e-3 < ev(s.reap_score - 1)


def test_evalue_formula():
    p, s = blz.search_setup(np.ones(500, np.uint8), 825_563_801, 2_000_000)
    for S in (40, 67, 100, 500):
        want = s.eff_searchsp * math.exp(-s.gapped.lambda_ * S + s.gapped.logK)
        assert blz.evalue(s, S) == want


def test_query_composition_statistics():
    """With query_composition the ungapped block is the query's letter mix against the background:
    a query drawn from the background gives nearly the ideal lambda, a tryptophan/cysteine-rich one
    a different block, each equal to the oracle's for that composition."""
    q_bg = blz.synth_query(5, 20000)
    rng = np.random.default_rng(3)
    q_wc = rng.choice(np.array([20, 3, 1, 11], np.uint8), 400, p=[0.4, 0.3, 0.2, 0.1]).astype(np.uint8)
    for q in (q_bg, q_wc):
        opts = blz.stats_options(query_composition=1)
        p, s = blz.search_setup(q, 10 ** 8, 10 ** 5, opts)
        freq = np.bincount(q, minlength=28).astype(float)
        lam, K, H = ostats.karlin(*ostats.score_distribution(freq))
        assert abs(s.ungapped.lambda_ - lam) < 1e-9 and abs(s.ungapped.K - K) < 1e-8
        assert p.x_ungapped == math.ceil(7 * math.log(2) / lam)
    ideal = blz.karlin_ungapped()
    _, s_bg = blz.search_setup(q_bg, 10 ** 8, 10 ** 5, blz.stats_options(query_composition=1))
    assert abs(s_bg.ungapped.lambda_ - ideal.lambda_) < 0.01


def test_hitlist_size_rule():
    """SBlastHitsParametersNew: composition-based statistics keep 2 x max_target_seqs + 50 lists,
    otherwise min(2x, x + 50)."""
    _, s = blz.search_setup(np.ones(100, np.uint8), 10 ** 6, 1000, blz.stats_options(max_target_seqs=500))
    assert s.hitlist_size == 1050
    _, s = blz.search_setup(np.ones(100, np.uint8), 10 ** 6, 1000,
                            blz.stats_options(max_target_seqs=30, composition_based_stats=0))
    assert s.hitlist_size == 60
