"""Capacity paths.  Device caps forced tiny through the tuning limits: a run whose work buffers
fill up runs again with larger ones, pairs past a per-subject cap go through the overflow queue
on the device, and every list still equals the oracle's.  In hand-back mode the same pairs are
reported (all of them, however short the overflow list) and every other list stays exact.

The reference's survivor buffer has no capacity argument at all (blast_engine.c:2965-2967,
gpuBLAZE.cu:1683); capped subjects are flagged 255 for the CPU (gpuBLAZE.cu:1088,1243-1247)."""
import os

import numpy as np
import pytest

import blz
import oracle

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
COUNTERS = ("word_hits", "ungapped_ext", "init_hits", "gapped_ext", "dp_cells")

TINY = dict(cap_stage=2, cap_retry=4, cap_ovf=6, cap_wide_ring=64, cap_wide_saved=3, cap_work=40,
            cap_pool=300, cap_hsps=60)


def _workload(seed=11, qlen=900, n=400):
    q = blz.synth_query(seed, qlen)
    spec = blz.synth_spec(seed=seed, n_subjects=n, len_median=800, len_max=20000, homolog_frac=0.25,
                          ident_min=0.3, ident_max=0.6, repeat_frac=0.05, indel_rate=0.01)
    res, offs, _ = blz.synth_db(spec, [q])
    return q, res, offs


def _check_lists(want, have, oids=None):
    keys = sorted(want) if oids is None else sorted(oids)
    assert sorted(have) == keys, (sorted(set(want) - set(have))[:5], sorted(set(have) - set(want))[:5])
    for oid in keys:
        assert np.array_equal(want[oid], have[oid]), oid


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("kw", [{}, dict(x_gapped=150)], ids=["default", "wide_bands"])
def test_tiny_caps_overflow_queue(variant, kw):
    """Every cap forced tiny: runs are repeated with grown buffers, the remaining pairs are
    finished by the overflow queue, and lists and work counters equal the oracle's.  With
    x_gapped 150 every band outgrows the wide pass's 64-entry ring, so every gapped subject goes
    through the queue."""
    q, res, offs = _workload()
    print(f"\n[case] variant {variant} params {kw}: {len(offs) - 1} subjects, {int(offs[-1])} residues",
          flush=True)
    with blz.Engine(0) as eng:
        eng.set_tuning(seed_variant=variant, **TINY)
        eng.upload_db(res, offs)
        h = eng.prepare(q, blz.default_params(**kw))
        try:
            got, st = eng.run(h)
        finally:
            eng.free_query(h)
    print(f"[case] reruns {st['reruns']}, queued pairs {st['overflow_subjects']}, rounds {st['drain_rounds']}, "
          f"queue {st['ms_drain']:.1f} ms", flush=True)
    counts, hsps, ost = oracle.prelim_db(q, res, offs, oracle.params(**kw), nthreads=NTHREADS)
    _check_lists(oracle.as_lists(counts, hsps), got.as_dict())
    for k in COUNTERS:
        assert st[k] == ost[k], k
    assert st["overflow_left"] == 0
    assert st["overflow_subjects"] > 0 and st["drain_rounds"] >= 1
    assert st["reruns"] >= 1


def test_tiny_caps_batch():
    """A batch of packed and unpacked queries under the same tiny caps: each query's lists equal
    the oracle's (pairs carry their own query through the queue)."""
    lens = [40, 120, 300, 650, 1400]
    qs = [blz.synth_query(60 + i, L) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=61, n_subjects=600, len_median=500, homolog_frac=0.2, ident_min=0.3,
                          ident_max=0.7, repeat_frac=0.05)
    res, offs, _ = blz.synth_db(spec, qs)
    print(f"\n[case] batch {lens} vs {len(offs) - 1} subjects", flush=True)
    with blz.Engine(0) as eng:
        eng.set_tuning(**TINY)
        eng.upload_db(res, offs)
        hs = [eng.prepare(q) for q in qs]
        try:
            outs, st = eng.run_batch(hs)
        finally:
            for h in hs:
                eng.free_query(h)
    assert st["overflow_left"] == 0 and st["overflow_subjects"] > 0
    for q, got in zip(qs, outs):
        counts, hsps, _ = oracle.prelim_db(q, res, offs, nthreads=NTHREADS)
        _check_lists(oracle.as_lists(counts, hsps), got.as_dict())


def test_hand_back_mode_reports_every_pair():
    """overflow_mode 1: buffers keep their (tiny) sizes, the run returns BLZ_EOVERFLOW, every
    pair past a cap is reported although the overflow list starts at 6 entries, no reported
    subject has a list, every other list equals the oracle's, and the reported pairs are exactly
    the oracle's subjects that are missing."""
    q, res, offs = _workload(seed=12)
    with blz.Engine(0) as eng:
        eng.set_tuning(overflow_mode=1, **TINY)
        eng.upload_db(res, offs)
        h = eng.prepare(q)
        try:
            got, st = eng.run(h, allow_overflow=True)
            pairs = eng.overflow_pairs()
        finally:
            eng.free_query(h)
    print(f"\n[case] {len(pairs)} pairs handed back, reruns {st['reruns']}", flush=True)
    assert st["overflow_left"] == len(pairs) > 6
    assert all(qi == 0 for qi, _ in pairs)
    handed = {s for _, s in pairs}
    assert len(handed) == len(pairs)
    counts, hsps, _ = oracle.prelim_db(q, res, offs, nthreads=NTHREADS)
    want = oracle.as_lists(counts, hsps)
    have = got.as_dict()
    assert not handed & set(have)
    _check_lists(want, have, oids=set(want) - handed)
    # the caller finishes the handed-back pairs on its own path: the oracle's lists for them
    # make the set whole
    assert set(have) | (handed & set(want)) == set(want)


def test_queue_drains_huge_init_hit_subject():
    """A 40,000-letter subject with 6,704 init hits against a 2,000-letter query (low cutoffs)
    is past the retry stage once cap_retry is 1,000: the queue stages it whole and its gapped
    pass keeps room for every init hit (1,799 HSPs survive the gapped cutoff)."""
    q = blz.synth_query(77, 2000)
    subs = [blz.synth_query(78, 40000), blz.synth_query(79, 300), blz.synth_query(80, 5000)]
    offs = np.zeros(len(subs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    res = np.concatenate(subs)
    kw = dict(ungapped_cutoff=12, gapped_cutoff=22, x_gapped=12)
    with blz.Engine(0) as eng:
        eng.set_tuning(cap_retry=1000, cap_wide_saved=3)
        eng.upload_db(res, offs)
        h = eng.prepare(q, blz.default_params(**kw))
        try:
            got, st = eng.run(h)
        finally:
            eng.free_query(h)
    counts, hsps, ost = oracle.prelim_db(q, res, offs, oracle.params(**kw), nthreads=NTHREADS)
    _check_lists(oracle.as_lists(counts, hsps), got.as_dict())
    for k in COUNTERS:
        assert st[k] == ost[k], k
    assert st["overflow_subjects"] >= 1 and st["overflow_left"] == 0
