"""The whole path end to end, as a caller runs a BLASTP search through the library: a makeblastdb
protein database (two volumes, formats 5 and 4, joined by an alias file, sequences in no length
order) read and made resident length-sorted; per query the search statistics (cutoffs, X-drops,
reap threshold); the prelim stage for a batch of queries; each query's hit list; the traceback of
every listed pair with the query's cutoff and the final X-drop; e-values.  Every step against the
oracle's on the database in OID order (the hit list's ties broken by OID, as NCBI's)."""
import os

import numpy as np
import pytest

import blz
import oracle
import stats as ostats
from test_blastdb import write_synthetic

pytestmark = pytest.mark.gpu

NTHREADS = len(os.sched_getaffinity(0))
SEARCH_FIELDS = ("word_threshold", "window", "x_ungapped", "x_gapped", "gap_open", "gap_extend", "ungapped_cutoff",
                 "gapped_cutoff")


def test_blastdb_to_traceback(tmp_path):
    lens = [90, 300, 640, 1500]
    queries = [blz.synth_query(4100 + i, L) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=41, n_subjects=12_000, homolog_frac=0.06, ident_min=0.3, ident_max=0.8,
                          sort_by_length=0)
    res, offs, _ = blz.synth_db(spec, queries)
    cut = 7000
    write_synthetic(tmp_path / "v1", res[: int(offs[cut])], offs[: cut + 1], version=5)
    write_synthetic(tmp_path / "v2", res[int(offs[cut]):], offs[cut:] - offs[cut], version=4)
    (tmp_path / "db.pal").write_text("TITLE pipeline\nDBLIST v1 v2\n")
    opts = blz.stats_options(expect=1.0, max_target_seqs=40, composition_based_stats=0)
    with blz.BlastDB(tmp_path / "db") as bdb, blz.Engine(0) as eng:
        perm = eng.upload_blastdb(bdb, sort_by_length=True)
        nres, nseq = bdb.num_residues, bdb.num_seqs
        setups = [blz.search_setup(q, nres, nseq, opts) for q in queries]
        hs = [eng.prepare(q, p) for q, (p, _) in zip(queries, setups)]
        try:
            eng.run_batch(hs, fetch=False)
            lists = [eng.hitlist(i, s)[0] for i, (_, s) in enumerate(setups)]
            tbp = blz.tb_params(x_dropoff=setups[0][1].x_final)
            got, st = eng.traceback(hs, lists, tbp, cutoffs=[s.reap_score for _, s in setups])
        finally:
            for h in hs:
                eng.free_query(h)
    assert nres == int(offs[-1]) and nseq == len(offs) - 1
    n_pairs = 0
    for qi, q in enumerate(queries):
        p, s = setups[qi]
        counts, hsps, _ = oracle.prelim_db(q, res, offs, oracle.params(**{f: getattr(p, f) for f in SEARCH_FIELDS}),
                                           nthreads=NTHREADS)
        want = ostats.reap(oracle.as_lists(counts, hsps), s.gapped.lambda_, s.gapped.K, s.eff_searchsp,
                           s.prelim_evalue)
        ranked = ostats.hitlist(want, s.gapped.lambda_, s.gapped.K, s.eff_searchsp, s.hitlist_size)
        positions = [int(x) for x in lists[qi].oids]
        assert [int(perm[x]) for x in positions] == [oid for oid, _ in ranked], qi
        for pos, (oid, ev) in zip(positions, ranked):
            subj = res[int(offs[oid]):int(offs[oid + 1])]
            assert np.array_equal(lists[qi].as_dict()[pos], want[oid])
            wt, wr = oracle.traceback_subject(q, subj, want[oid], x_final=tbp.x_dropoff, cutoff=s.reap_score)
            gt, gr = got[qi][pos]
            assert gt.tolist() == wt.tolist() and gr.tolist() == wr.tolist(), (qi, oid)
            for rec in gt:
                e = blz.evalue(s, int(rec["score"]))
                assert e <= s.prelim_evalue
                assert e == pytest.approx(ostats.evalue(int(rec["score"]), s.gapped.lambda_, s.gapped.K,
                                                        s.eff_searchsp), rel=1e-12)
            n_pairs += 1
    print(f"\n[case] {len(queries)} queries, {nseq} subjects in two volumes: {n_pairs} listed pairs traced, "
          f"{st['hsps']} HSPs, all equal to the oracle", flush=True)
    assert n_pairs > 40
