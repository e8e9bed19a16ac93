"""A database read from makeblastdb-format files, made resident in length order: after mapping
positions back to OIDs, every list equals the oracle's over the database in OID order."""
import os

import numpy as np
import pytest

import blz
import oracle
from test_blastdb import write_synthetic

pytestmark = pytest.mark.gpu

NTHREADS = len(os.sched_getaffinity(0))


@pytest.mark.parametrize("version", [4, 5])
def test_blastdb_upload_sorted(tmp_path, version):
    q = blz.synth_query(8, 400)
    spec = blz.synth_spec(seed=8, n_subjects=20_000, homolog_frac=0.05, ident_min=0.35, ident_max=0.8,
                          sort_by_length=0)
    res, offs, _ = blz.synth_db(spec, [q])
    write_synthetic(tmp_path / "db", res, offs, version)
    with blz.BlastDB(tmp_path / "db") as bdb, blz.Engine(0) as eng:
        perm = eng.upload_blastdb(bdb, sort_by_length=True)
        h = eng.prepare(q)
        got, _ = eng.run(h)
        eng.free_query(h)
    have = {int(perm[k]): v for k, v in got.as_dict().items()}
    counts, hsps, _ = oracle.prelim_db(q, res, offs, nthreads=NTHREADS)
    want = oracle.as_lists(counts, hsps)
    print(f"\n[case] version {version}: {len(want)} subjects with lists", flush=True)
    assert sorted(want) == sorted(have) and len(want) > 100
    assert all(np.array_equal(want[k], have[k]) for k in want)
