"""Prelim e-value reap on the device and the per-query hit list, against the oracle.

The library's search setup gives each query its cutoffs and reap score; the device drops the
HSPs above the prelim e-value after the purge and sort (blast_engine.c:1436-1457,896-930), and
blz_result_hitlist orders the subjects as NCBI's hit list does (prelim_stage.cpp:1462-1498).
The oracle runs the prelim stage with the same parameters, then reaps by e-value and ranks with
its own restatement (oracle/stats.py)."""
import os

import numpy as np
import pytest

import blz
import oracle
import stats as ostats

pytestmark = pytest.mark.gpu

NTHREADS = len(os.sched_getaffinity(0))
SEARCH_FIELDS = ("word_threshold", "window", "x_ungapped", "x_gapped", "gap_open", "gap_extend", "ungapped_cutoff",
                 "gapped_cutoff")


def oracle_lists(q, res, offs, p):
    counts, hsps, _ = oracle.prelim_db(q, res, offs, oracle.params(**{f: getattr(p, f) for f in SEARCH_FIELDS}),
                                       nthreads=NTHREADS)
    return oracle.as_lists(counts, hsps)


def check_query(eng, qi, q, res, offs, p, s):
    got = eng._fetch(qi).as_dict()
    want_all = oracle_lists(q, res, offs, p)
    want = ostats.reap(want_all, s.gapped.lambda_, s.gapped.K, s.eff_searchsp, s.prelim_evalue)
    assert sorted(want) == sorted(got)
    assert all(np.array_equal(want[k], got[k]) for k in want)
    reaped = sum(len(v) for v in want_all.values()) - sum(len(v) for v in want.values())
    hl, ev = eng.hitlist(qi, s)
    ranked = ostats.hitlist(want, s.gapped.lambda_, s.gapped.K, s.eff_searchsp, s.hitlist_size)
    assert [int(x) for x in hl.oids] == [oid for oid, _ in ranked]
    assert np.allclose(ev, [e for _, e in ranked], rtol=1e-12, atol=0)
    d = hl.as_dict()
    assert all(np.array_equal(d[k], want[k]) for k in d)
    return reaped, len(hl.oids)


def test_reap_and_hitlist():
    q = blz.synth_query(21, 300)
    spec = blz.synth_spec(seed=21, n_subjects=20_000, homolog_frac=0.08, ident_min=0.3, ident_max=0.8)
    res, offs, _ = blz.synth_db(spec, [q])
    opts = blz.stats_options(prelim_evalue=1e-4, max_target_seqs=25, composition_based_stats=0)
    p, s = blz.search_setup(q, int(offs[-1]), len(offs) - 1, opts)
    assert p.reap_score > p.gapped_cutoff
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        h = eng.prepare(q, p)
        eng.run(h)
        reaped, listed = check_query(eng, 0, q, res, offs, p, s)
        eng.free_query(h)
    print(f"\n[case] reap score {p.reap_score} (cutoff {p.gapped_cutoff}): {reaped} HSPs reaped, {listed} listed",
          flush=True)
    assert reaped > 0 and listed == 50


def test_reap_per_query_in_a_packed_batch():
    """Short queries of different lengths share seed scans while each keeps its own gapped cutoff
    and reap score from the search setup."""
    lens = [40, 90, 150, 200, 260, 320, 700]
    qs = [blz.synth_query(400 + i, L) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=44, n_subjects=8000, homolog_frac=0.1, ident_min=0.3, ident_max=0.8)
    res, offs, _ = blz.synth_db(spec, qs)
    opts = blz.stats_options(prelim_evalue=1e-3, max_target_seqs=10)
    setups = [blz.search_setup(q, int(offs[-1]), len(offs) - 1, opts) for q in qs]
    assert len({p.gapped_cutoff for p, _ in setups}) > 1
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        hs = [eng.prepare(q, p) for q, (p, _) in zip(qs, setups)]
        try:
            _, st = eng.run_batch(hs, fetch=False)
            for i, (q, (p, s)) in enumerate(zip(qs, setups)):
                check_query(eng, i, q, res, offs, p, s)
        finally:
            for h in hs:
                eng.free_query(h)
    assert st["seed_launches"] >= 1
