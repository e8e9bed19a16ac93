"""blz_prelim_run_noalloc (every work buffer in the library's fixed per-GPU array) through the C ABI:
its lists equal blz_prelim_run's on two calls in a row, the second on a smaller database and query
(so stale contents of the shared array from the first call would show), it rejects a database past
its subject limit, and reports a too-small sink with the sizes."""
import ctypes as C

import numpy as np
import pytest

import blz

pytestmark = pytest.mark.gpu


def run_noalloc(eng, qh, cap_subjects, cap_hsps):
    lib = eng.lib
    sink, (oids, counts, hsps) = blz.Engine._sink(cap_subjects, cap_hsps)
    stats = blz.RunStats()
    st = lib.blz_prelim_run_noalloc(eng.ctx, qh, eng.db, C.byref(sink), C.byref(stats))
    return st, sink, blz.Result(oids[: sink.n_subjects], counts[: sink.n_subjects], hsps[: sink.n_hsps]), stats


def test_noalloc_matches_prelim_run_twice():
    with blz.Engine(0) as eng:
        for call, (seed, qlen, nsubj) in enumerate([(71, 300, 20_000), (72, 120, 5_000)]):
            q = blz.synth_query(seed, qlen)
            spec = blz.synth_spec(seed=seed, n_subjects=nsubj, homolog_frac=0.05, ident_min=0.35, ident_max=0.8)
            res, offs, _ = blz.synth_db(spec, [q])
            eng.upload_db(res, offs)
            qh = eng.prepare(q)
            try:
                want, wst = eng.run(qh)
                st, sink, got, gst = run_noalloc(eng, qh, len(offs) - 1, 1 << 18)
            finally:
                eng.free_query(qh)
            print(f"\n[case] call {call}: {nsubj} subjects, query {qlen} aa: {sink.n_subjects} lists, "
                  f"{sink.n_hsps} HSPs, status {st}", flush=True)
            assert st == 0, eng.lib.blz_last_error(eng.ctx)
            assert np.array_equal(got.oids, want.oids)
            assert np.array_equal(got.counts, want.counts)
            assert np.array_equal(got.hsps, want.hsps)
            for k in ("word_hits", "ungapped_ext", "init_hits", "gapped_ext", "dp_cells", "hsps"):
                assert getattr(gst, k) == wst[k], k
            assert len(want.oids) > 10


def test_noalloc_limits():
    with blz.Engine(0) as eng:
        q = blz.synth_query(73, 200)
        spec = blz.synth_spec(seed=73, n_subjects=3000, homolog_frac=0.1, ident_min=0.4, ident_max=0.8)
        res, offs, _ = blz.synth_db(spec, [q])
        eng.upload_db(res, offs)
        qh = eng.prepare(q)
        try:
            want, _ = eng.run(qh)
            st, sink, _, _ = run_noalloc(eng, qh, len(want.oids), max(0, len(want.hsps) - 1))
            assert st == 4 and sink.n_subjects == len(want.oids) and sink.n_hsps == len(want.hsps)
            big_offs = np.arange(0, 5 * 131_073 + 1, 5, dtype=np.uint64)
            eng.upload_db(np.ones(int(big_offs[-1]), np.uint8), big_offs)
            st, _, _, _ = run_noalloc(eng, qh, 16, 16)
            assert st == 1
            assert b"131072" in eng.lib.blz_last_error(eng.ctx)
        finally:
            eng.free_query(qh)
