"""Serving loop (blz_server_*): continuous batching over a resident database.  Each request's
lists must be exactly what the query gets alone (blz_prelim_run), whatever batch it ran in."""
import os
import threading
import time

import numpy as np
import pytest

import blz
import oracle

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


def test_serving_alone_vs_mix():
    rng = np.random.default_rng(3)
    lens = [int(x) for x in np.clip(np.exp(rng.normal(np.log(350), 0.8, 48)), 30, 2500)]
    qs = [blz.synth_query(900 + i, L) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=23, n_subjects=3000, homolog_frac=0.05, ident_min=0.35, ident_max=0.8)
    res, offs, _ = blz.synth_db(spec, qs)
    print(f"\n[case] {len(qs)} queries ({sum(lens)} letters) vs {len(offs) - 1} subjects", flush=True)
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        alone = []
        for q in qs:
            h = eng.prepare(q)
            r, _ = eng.run(h)
            eng.free_query(h)
            alone.append(r)
        with blz.Server(eng, slots=3, max_batch=8, max_batch_letters=4000, linger_us=20000) as srv:
            tickets = [None] * len(qs)

            def submitter(k):  # four client threads, staggered arrivals
                for i in range(k, len(qs), 4):
                    tickets[i] = srv.submit(qs[i])
                    time.sleep(0.0005 * (i % 3))

            ths = [threading.Thread(target=submitter, args=(k,)) for k in range(4)]
            for t in ths:
                t.start()
            for t in ths:
                t.join()
            sizes = []
            for i, tk in enumerate(tickets):
                info = srv.wait(tk)
                assert info["status"] == 0, info
                got = srv.result(tk)
                srv.release(tk)
                sizes.append(info["batch_queries"])
                assert np.array_equal(got.oids, alone[i].oids), i
                assert np.array_equal(got.counts, alone[i].counts), i
                assert np.array_equal(got.hsps, alone[i].hsps), i
            st = srv.stats()
            print(f"[case] {st['batches']} batches, batch sizes seen {sorted(set(sizes))}", flush=True)
            assert st["completed"] == len(qs)
            assert max(sizes) > 1  # requests were batched
    for i in (0, len(qs) // 2):  # and alone == the CPU restatement
        counts, hsps, _ = oracle.prelim_db(qs[i], res, offs, nthreads=NTHREADS)
        want = oracle.as_lists(counts, hsps)
        have = alone[i].as_dict()
        assert sorted(want) == sorted(have)
        assert all(np.array_equal(want[k], have[k]) for k in want)


def _requests(seed, n, db_subjects):
    rng = np.random.default_rng(seed)
    lens = [int(x) for x in np.clip(np.exp(rng.normal(np.log(350), 0.8, n)), 30, 2500)]
    qs = [blz.synth_query(seed * 1000 + i, L) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=seed, n_subjects=db_subjects, homolog_frac=0.1, ident_min=0.3, ident_max=0.8,
                          repeat_frac=0.05)
    res, offs, _ = blz.synth_db(spec, qs)
    return qs, res, offs


def test_admission_by_footprint():
    """A footprint budget of one work item admits one request per batch although max_batch is 8;
    a generous budget batches them; the lists are the solo runs' either way."""
    qs, res, offs = _requests(31, 16, 3000)
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        alone = []
        for q in qs:
            h = eng.prepare(q)
            alone.append(eng.run(h)[0])
            eng.free_query(h)
        for budget, want_single in ((1, True), (10 ** 9, False)):
            with blz.Server(eng, slots=2, max_batch=8, max_batch_work=budget, linger_us=20000) as srv:
                tickets = [srv.submit(q) for q in qs]
                sizes = []
                for i, tk in enumerate(tickets):
                    info = srv.wait(tk)
                    assert info["status"] == 0
                    got = srv.result(tk)
                    srv.release(tk)
                    sizes.append(info["batch_queries"])
                    assert np.array_equal(got.hsps, alone[i].hsps) and np.array_equal(got.oids, alone[i].oids)
                st = srv.stats()
            print(f"\n[case] budget {budget}: batch sizes {sorted(set(sizes))}, rate {st['work_rate']:.3g}", flush=True)
            assert (max(sizes) == 1) == want_single
            assert st["work_rate"] > 0


def test_served_overflow_subjects_per_ticket():
    """Hand-back mode with tiny caps: each request reports the subjects its batch could not finish
    (blz_server_overflow); its lists equal the oracle's for every other subject, and the oracle's
    subjects missing from them are among the reported ones."""
    qs, res, offs = _requests(32, 6, 1500)
    tune = dict(overflow_mode=1, cap_stage=2, cap_retry=4, cap_ovf=8, cap_wide_ring=64, cap_wide_saved=3)
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        with blz.Server(eng, slots=1, max_batch=6, tuning=tune) as srv:
            tickets = [srv.submit(q) for q in qs]
            out = []
            for tk in tickets:
                info = srv.wait(tk)
                assert info["status"] in (0, 5)
                got = srv.result(tk)
                out.append((got, set(srv.overflow(tk)), info))
                srv.release(tk)
    assert sum(len(o) for _, o, _ in out) > 0
    for i in (0, 3):
        got, ovf, info = out[i]
        assert info["n_overflow"] == len(ovf)
        counts, hsps, _ = oracle.prelim_db(qs[i], res, offs, nthreads=NTHREADS)
        want = oracle.as_lists(counts, hsps)
        have = got.as_dict()
        assert not ovf & set(have)
        assert set(want) - set(have) <= ovf
        for k in have:
            assert np.array_equal(want[k], have[k]), (i, k)
