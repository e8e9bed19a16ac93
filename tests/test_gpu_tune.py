"""The autotuner on the device: it times real batched runs over every knob it searches, ends on
a choice no slower than the built-in one, and the lists under that choice equal the built-in
choice's, query for query."""
import numpy as np
import pytest

import blz
import blz_tune

pytestmark = pytest.mark.gpu


def test_tuned_choice_keeps_lists():
    rng = np.random.default_rng(51)
    lens = np.exp(rng.uniform(np.log(50), np.log(1500), 12)).astype(int)
    queries = [blz.synth_query(5100 + i, int(L)) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=51, n_subjects=30_000, homolog_frac=0.05, ident_min=0.35, ident_max=0.7)
    res, offs, _ = blz.synth_db(spec, queries)
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        hs = [eng.prepare(q) for q in queries]
        try:
            base, _ = eng.run_batch(hs)
            base_ms = blz_tune.time_batch(eng, hs, {**eng.get_tuning(), "graph": 1}, repeats=2)
            lines = []
            best, ms = blz_tune.tune(eng, hs, repeats=2, log=lines.append)
            tuned, _ = eng.run_batch(hs)
        finally:
            for h in hs:
                eng.free_query(h)
    print(f"\n[case] built-in {base_ms:.2f} ms, tuned {ms:.2f} ms: "
          f"{ {k: v for k, v in best.items() if k in blz_tune.CANDIDATES and v} }", flush=True)
    searched = {line.split()[1].split("=")[0] for line in lines if "=" in line}
    assert searched == set(blz_tune.CANDIDATES)
    assert ms <= base_ms * 1.10
    for a, b in zip(base, tuned):
        assert np.array_equal(a.oids, b.oids) and np.array_equal(a.counts, b.counts) and np.array_equal(a.hsps, b.hsps)
