"""Each benchmark workload on the GPU at its stated size, checked against the oracle.

config 0: 1 x 300 aa vs 10,000 subjects (the CPU reference case), every list against the oracle;
  the reference's own 40 benchmark queries (BLAZE_queries/*.fa, NP_937947 of 7,095 aa included)
  against the same database with ambiguity letters planted in it, as one batch.
config 2: 256 queries (log-uniform 50-3,000 aa) vs the 2,000,000-subject database: the batched
  run equals 256 plain runs list for list, and two queries equal the oracle.
config 3: 64 queries of 1,000-2,000 aa vs 500,000 long subjects with repeats: two queries equal
  the oracle (lists and work counters); with the per-subject caps forced small, the overflow
  queue takes thousands of pairs and every list of the batch is unchanged.
config 4: a 40,000,000-subject database resident in HBM, requests of up to 4,000 aa through the
  serving loop: each request's lists equal its solo run, and one equals the oracle.
"""
import os
import time

import numpy as np
import pytest

import blz
import oracle

pytestmark = pytest.mark.gpu

NTHREADS = len(os.sched_getaffinity(0))
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "queries_ncbistdaa.txt")
AMBIGUOUS = np.array([2, 21, 23, 24, 25, 26, 27], np.uint8)  # B X Z U * O J
COUNTERS = ("word_hits", "ungapped_ext", "init_hits", "gapped_ext", "dp_cells")


def log(msg):
    print(f"[case] {msg}", flush=True)


def check(q, res, offs, got, p=None, stats=None):
    counts, hsps, ost = oracle.prelim_db(q, res, offs, p, nthreads=NTHREADS)
    want = oracle.as_lists(counts, hsps)
    have = got.as_dict()
    assert sorted(want) == sorted(have), (sorted(set(want) ^ set(have))[:10])
    bad = [k for k in want if not np.array_equal(want[k], have[k])]
    assert not bad, f"{len(bad)} lists differ, first subject {bad[0]}"
    if stats is not None:
        for k in COUNTERS:
            assert stats[k] == ost[k], k
    return len(want)


def same(a, b):
    return np.array_equal(a.oids, b.oids) and np.array_equal(a.counts, b.counts) and np.array_equal(a.hsps, b.hsps)


def golden_queries():
    out = []
    for line in open(GOLDEN):
        name, *codes = line.split()
        out.append((name, np.array(list(map(int, codes)), np.uint8)))
    return out


def test_config0_full_size():
    q = blz.synth_query(0, 300)
    spec = blz.synth_spec(seed=0, n_subjects=10_000)  # the defaults are config 0's
    res, offs, hom = blz.synth_db(spec, [q])
    log(f"config 0: {len(offs) - 1} subjects, {int(offs[-1])} residues")
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        h = eng.prepare(q)
        got, st = eng.run(h)
        eng.free_query(h)
    n = check(q, res, offs, got, stats=st)
    assert n >= 0.9 * (hom >= 0).sum()


def test_reference_queries_with_ambiguity_letters():
    """The reference's 40 benchmark queries (36-7,095 aa; they contain no ambiguity letter) plus
    three of them with 2% of their letters replaced by B/X/Z/U/*/O/J, against the config-0
    database with 1% of its residues replaced by those letters; one batched run, every query's
    lists against the oracle."""
    rng = np.random.default_rng(9)
    named = golden_queries()
    queries = [q for _, q in named]
    assert max(len(q) for q in queries) == 7095
    for k in (0, 7, 21):
        v = queries[k].copy()
        pos = rng.random(len(v)) < 0.02
        v[pos] = rng.choice(AMBIGUOUS, int(pos.sum()))
        queries.append(v)
    spec = blz.synth_spec(seed=0, n_subjects=10_000, homolog_frac=0.05)
    res, offs, _ = blz.synth_db(spec, queries)
    res = res.copy()
    pos = rng.random(res.size) < 0.01
    res[pos] = rng.choice(AMBIGUOUS, int(pos.sum()))
    log(f"{len(queries)} queries ({sum(map(len, queries))} letters) vs {len(offs) - 1} subjects, "
        f"{int(pos.sum())} ambiguity letters planted")
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        hs = [eng.prepare(q) for q in queries]
        try:
            outs, st = eng.run_batch(hs)
        finally:
            for h in hs:
                eng.free_query(h)
    total = 0
    for q, got in zip(queries, outs):
        total += check(q, res, offs, got)
    log(f"{total} subject lists checked")
    assert total > 0


@pytest.fixture(scope="module")
def db2m():
    spec = blz.synth_spec(seed=0, n_subjects=2_000_000, homolog_frac=0.01, ident_min=0.35, ident_max=0.70)
    q = blz.synth_query(0, 500)
    t = time.perf_counter()
    res, offs, _ = blz.synth_db(spec, [q])
    log(f"2M-subject DB built in {time.perf_counter() - t:.1f} s: {int(offs[-1])} residues")
    return res, offs


def test_config2_batch_equals_plain_and_oracle(db2m):
    res, offs = db2m
    rng = np.random.default_rng(0)
    lens = np.exp(rng.uniform(np.log(50), np.log(3000), 256)).astype(int)
    queries = [blz.synth_query(1000 + i, int(L)) for i, L in enumerate(lens)]
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        hs = [eng.prepare(q) for q in queries]
        try:
            t = time.perf_counter()
            outs, st = eng.run_batch(hs)
            log(f"batch of 256 ({int(lens.sum())} letters): {time.perf_counter() - t:.2f} s, {st['seed_launches']} seed launches")
            assert st["overflow_left"] == 0
            tune = eng.get_tuning()
            eng.set_tuning(**{**tune, "query_packing": 1, "graph": 1})
            t = time.perf_counter()
            for i, h in enumerate(hs):
                r, _ = eng.run(h)
                assert same(r, outs[i]), (i, int(lens[i]))
            log(f"256 plain runs: {time.perf_counter() - t:.2f} s, all lists equal")
        finally:
            for h in hs:
                eng.free_query(h)
    for i in (int(np.argmin(lens)), int(np.argmax(lens))):
        log(f"oracle: query {i} ({int(lens[i])} aa)")
        check(queries[i], res, offs, outs[i])


def test_config3_full_size_and_overflow_queue():
    rng = np.random.default_rng(0)
    lens = rng.integers(1000, 2001, 64)
    queries = [blz.synth_query(2000 + i, int(L)) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=0, n_subjects=500_000, len_median=800, len_sigma=0.8, len_min=30, len_max=35000,
                          homolog_frac=0.2, ident_min=0.30, ident_max=0.45, repeat_frac=0.05)
    t = time.perf_counter()
    res, offs, _ = blz.synth_db(spec, queries)
    log(f"config 3 DB built in {time.perf_counter() - t:.1f} s: {len(offs) - 1} subjects, {int(offs[-1])} residues")
    sample = [0, 32]
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        hs = [eng.prepare(q) for q in queries]
        try:
            t = time.perf_counter()
            outs, st = eng.run_batch(hs)
            log(f"batch: {time.perf_counter() - t:.2f} s, {st['dp_cells']} DP cells, {st['overflow_subjects']} queued pairs")
            alone = {}
            for i in sample:
                alone[i] = eng.run(hs[i])
                assert same(alone[i][0], outs[i])
            # per-subject caps forced small: the overflow queue finishes what the passes cannot
            tune = eng.get_tuning()
            eng.set_tuning(**{**tune, "cap_stage": 4, "cap_retry": 8, "cap_wide_ring": 128, "cap_wide_saved": 4})
            t = time.perf_counter()
            capped, cst = eng.run_batch(hs)
            log(f"capped batch: {time.perf_counter() - t:.2f} s, {cst['overflow_subjects']} pairs through the overflow "
                f"queue in {cst['drain_rounds']} rounds ({cst['ms_drain']:.0f} ms)")
            assert cst["overflow_subjects"] > 1000 and cst["overflow_left"] == 0
            for i in range(len(queries)):
                assert same(capped[i], outs[i]), i
            for k in COUNTERS:
                assert cst[k] == st[k], k
        finally:
            for h in hs:
                eng.free_query(h)
    for i in sample:
        log(f"oracle: query {i} ({int(lens[i])} aa)")
        check(queries[i], res, offs, outs[i], stats=alone[i][1])


def test_config4_served_40m_subjects():
    rng = np.random.default_rng(4)
    lens = [int(x) for x in np.clip(np.exp(rng.normal(np.log(350), 0.8, 10)), 30, 4000)] + [4000, 3500]
    queries = [blz.synth_query(5000 + i, L) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=0, n_subjects=40_000_000, homolog_frac=0.01, ident_min=0.35, ident_max=0.70)
    t = time.perf_counter()
    res, offs, _ = blz.synth_db(spec, queries[:4])
    log(f"config 4 DB built in {time.perf_counter() - t:.1f} s: {len(offs) - 1} subjects, {int(offs[-1])} residues")
    with blz.Engine(0) as eng:
        t = time.perf_counter()
        eng.upload_db(res, offs)
        log(f"resident in HBM after {time.perf_counter() - t:.1f} s")
        alone = []
        for q in queries:
            h = eng.prepare(q)
            r, _ = eng.run(h)
            eng.free_query(h)
            alone.append(r)
        log("solo runs done")
        with blz.Server(eng, slots=3, max_batch=8) as srv:
            tickets = [srv.submit(q) for q in queries]
            sizes = []
            for i, tk in enumerate(tickets):
                info = srv.wait(tk)
                assert info["status"] == 0, info
                got = srv.result(tk)
                srv.release(tk)
                sizes.append(info["batch_queries"])
                assert same(got, alone[i]), (i, lens[i])
        log(f"served: batch sizes {sizes}")
        assert max(sizes) > 1
    i = int(np.argmin(np.abs(np.array(lens) - 350)))
    log(f"oracle: request {i} ({lens[i]} aa) over {int(offs[-1])} residues")
    check(queries[i], res, offs, alone[i])
