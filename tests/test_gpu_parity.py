"""GPU vs oracle: per-subject HSP lists must be identical (bit-exact: same subjects, same HSPs,
same order, identical offsets, gapped starts and raw scores)."""
import os

import numpy as np
import pytest

import blz
import oracle

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


# seed kernels: 1 = global tables with in-place extension, 2 = the batched kernel (short
# subjects of a single-query scan share a wave's scan), 3 = the batched kernel scanning one
# subject at a time; the first also runs the gapped DP with whole waves per subject, the others
# with four subjects per wave, the third with the packed 16-bit DP (two columns per lane)
@pytest.fixture(scope="module", params=[1, 2, 3], ids=["seed_v1", "seed_v2", "seed_v2_single"])
def engine(request):
    print(f"\n[engine] seed variant {request.param}", flush=True)
    with blz.Engine(0) as eng:
        eng.set_tuning(seed_variant=min(request.param, 2), gapped_lanes=64 if request.param == 1 else 16,
                       subject_groups=1 if request.param == 3 else 0, gapped_width=16 if request.param == 3 else 0)
        yield eng


_ORACLE_CACHE = {}


def oracle_run(query, res, offs, p_ora):
    key = (query.tobytes(), res.size, int(offs[-1]), hash(res[: 1 << 16].tobytes()), len(offs),
           None if p_ora is None else tuple(getattr(p_ora, f) for f, _ in p_ora._fields_))
    if key not in _ORACLE_CACHE:
        _ORACLE_CACHE.clear()
        _ORACLE_CACHE[key] = oracle.prelim_db(query, res, offs, p_ora, nthreads=NTHREADS)
    return _ORACLE_CACHE[key]


def compare(eng, query, res, offs, p_blz=None, p_ora=None, allow_overflow=False):
    print(f"[case] query {len(query)} aa, {len(offs) - 1} subjects, {int(offs[-1])} residues", flush=True)
    eng.upload_db(res, offs)
    qh = eng.prepare(query, p_blz)
    try:
        got, stats = eng.run(qh, allow_overflow=allow_overflow)
    finally:
        eng.free_query(qh)
    counts, hsps, ost = oracle_run(query, res, offs, p_ora)
    want = oracle.as_lists(counts, hsps)
    have = got.as_dict()
    missing = sorted(set(want) - set(have))
    extra = sorted(set(have) - set(want))
    assert not missing and not extra, f"subjects differ: missing {missing[:10]} extra {extra[:10]}"
    bad = [oid for oid in want if not np.array_equal(want[oid], have[oid])]
    assert not bad, f"{len(bad)} subjects differ, first {bad[0]}: want {want[bad[0]]} got {have[bad[0]]}"
    # the device counters agree with the oracle's on the work both did
    assert stats["word_hits"] == ost["word_hits"]
    assert stats["ungapped_ext"] == ost["ungapped_ext"]
    assert stats["init_hits"] == ost["init_hits"]
    assert stats["gapped_ext"] == ost["gapped_ext"]
    assert stats["dp_cells"] == ost["dp_cells"]
    return stats, len(want)


def make(seed, nq_len, n, **kw):
    q = blz.synth_query(seed, nq_len)
    spec = blz.synth_spec(seed=seed, n_subjects=n, **kw)
    res, offs, hom = blz.synth_db(spec, [q])
    return q, res, offs, hom


def test_config0_cut(engine):
    """Config 0 shape (300 aa query, log-normal median 300, 2% homologs at 50%), 3000 subjects."""
    q, res, offs, hom = make(0, 300, 3000, homolog_frac=0.02, ident_min=0.5, ident_max=0.5)
    stats, nsub = compare(engine, q, res, offs)
    assert nsub >= 0.9 * (hom >= 0).sum()
    assert stats["overflow_subjects"] == 0


@pytest.mark.parametrize("qlen", [3, 17, 64, 500, 1000, 2000])
def test_query_lengths(engine, qlen):
    q, res, offs, _ = make(7 + qlen, qlen, 800, homolog_frac=0.05, ident_min=0.35, ident_max=0.7)
    compare(engine, q, res, offs)


@pytest.mark.parametrize("case", range(6))
def test_random_workloads(engine, case):
    """Workloads drawn at random: query length log-uniform 20-2,500 aa, 60,000 subjects of a
    drawn length distribution (sorted or not), homolog fraction, identity range, indels and
    repeat blocks; every list and counter matches the oracle."""
    rng = np.random.default_rng(4100 + case)
    qlen = int(np.exp(rng.uniform(np.log(20), np.log(2500))))
    ident_min = float(rng.uniform(0.3, 0.7))
    kw = dict(homolog_frac=float(rng.choice([0.0, 0.01, 0.05, 0.2])), ident_min=ident_min,
              ident_max=min(0.95, ident_min + float(rng.uniform(0.0, 0.3))),
              len_median=float(rng.choice([60.0, 300.0, 900.0])), len_sigma=float(rng.uniform(0.3, 1.0)),
              indel_rate=float(rng.choice([0.0, 0.02])), repeat_frac=float(rng.choice([0.0, 0.05])),
              sort_by_length=int(rng.integers(0, 2)))
    print(f"[case] random {case}: qlen {qlen} {kw}", flush=True)
    q, res, offs, _ = make(4200 + case, qlen, 60000, **kw)
    compare(engine, q, res, offs)


@pytest.mark.parametrize("case", range(6))
def test_random_params(engine, case):
    """Search parameters drawn at random (word threshold, two-hit window, both X-drops, gap costs,
    both cutoffs) over a homolog-rich workload; every list and counter matches the oracle."""
    rng = np.random.default_rng(4300 + case)
    kw = dict(word_threshold=int(rng.integers(10, 14)), window=int(rng.integers(8, 120)),
              x_ungapped=int(rng.integers(6, 48)), x_gapped=int(rng.integers(10, 80)), gap_open=int(rng.integers(3, 16)),
              gap_extend=int(rng.integers(1, 4)), ungapped_cutoff=int(rng.integers(18, 70)),
              gapped_cutoff=int(rng.integers(25, 90)))
    print(f"[case] random params {case}: {kw}", flush=True)
    q, res, offs, _ = make(4400 + case, int(rng.integers(100, 1500)), 30000, homolog_frac=0.05,
                           ident_min=0.3, ident_max=0.8, indel_rate=0.02)
    compare(engine, q, res, offs, blz.default_params(**kw), oracle.params(**kw))


@pytest.mark.parametrize("window", [1, 2, 3, 4])
def test_small_windows(engine, window):
    """Two-hit windows at and below the word size.  With a window of 3 or less no pair is ever
    extended (a second hit is either an overlap or its diagonal's new last hit), and hits that lose
    a claim and are settled against their diagonal's new state follow the same rule; every list
    and counter matches the oracle."""
    kw = dict(window=window, ungapped_cutoff=20)
    print(f"[case] small window {window}", flush=True)
    q, res, offs, _ = make(4500 + window, 400, 20000, homolog_frac=0.05, ident_min=0.4, ident_max=0.9)
    compare(engine, q, res, offs, blz.default_params(**kw), oracle.params(**kw))


def test_edge_subjects(engine):
    """Subjects of length 1, 2, 3, the query itself, a reversed query and empty-range neighbours."""
    q = blz.synth_query(3, 200)
    subs = [q[:1], q[:2], q[:3], q.copy(), q[::-1].copy(), q[50:120].copy(), np.full(40, 1, np.uint8),
            q[:0], np.concatenate([q, q]).copy()]
    offs = np.zeros(len(subs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    res = np.concatenate(subs).astype(np.uint8)
    compare(engine, q, res, offs)


def _mutated(rng, src, ident):
    """Copy of src with a fraction 1 - ident of its letters replaced by random standard letters."""
    out = src.copy()
    std = np.array([1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22], np.uint8)
    m = rng.random(len(out)) > ident
    out[m] = std[rng.integers(0, len(std), size=int(m.sum()))]
    return out


def test_subject_group_boundaries(engine):
    """Subjects around the sizes where a wave's group of subjects closes: eight-subject groups of
    tiny subjects, 1,020-1,025 letters (alone, or too long for the LDS buffer), pairs whose
    64-aligned starts just fit or just miss 1,024 letters, with homologous stretches at subject
    starts and ends so HSPs and extensions run up to the zero separators."""
    rng = np.random.default_rng(77)
    q = blz.synth_query(31, 600)
    lens = [3, 5, 9, 17, 30, 31, 40, 2, 4, 64, 60, 61, 1020, 1021, 1024, 1025, 500, 520, 448, 508,
            459, 505, 100, 900, 200, 200, 200, 200, 200, 12, 700, 300]
    subs = []
    for k, L in enumerate(lens):
        s = blz.synth_query(900 + k, L).copy()
        if L >= 60:  # homologous stretch of the query at the start and at the end
            n = min(L // 2, 150)
            s[:n] = _mutated(rng, q[100:100 + n], 0.7)
            s[L - n:] = _mutated(rng, q[300:300 + n], 0.8)
        subs.append(s)
    offs = np.zeros(len(subs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    stats, nsub = compare(engine, q, np.concatenate(subs), offs)
    assert nsub >= 10


def test_wide_window_and_ungapped_xdrop_cap(engine):
    """A two-hit window above 1,024 (subject groups off) is honoured; an ungapped X-drop above
    16,384 (walks stop on letter 0's -32,768) is refused."""
    q, res, offs, _ = make(23, 300, 800, homolog_frac=0.1, ident_min=0.4, ident_max=0.6)
    kw = dict(window=1500)
    compare(engine, q, res, offs, blz.default_params(**kw), oracle.params(**kw))
    with pytest.raises(blz.BlzError) as e:
        engine.prepare(q, blz.default_params(x_ungapped=16385))
    assert e.value.status == 1  # BLZ_EINVAL


def test_low_complexity(engine):
    """Poly-residue and short-period repeats: many hits per word and per diagonal."""
    q = np.concatenate([np.full(120, 1, np.uint8), np.tile(np.array([11, 4, 9], np.uint8), 60),
                        blz.synth_query(5, 100)])
    subs = [np.full(300, 1, np.uint8), np.tile(np.array([11, 4, 9], np.uint8), 100),
            np.tile(np.array([11, 4, 9, 9], np.uint8), 70), blz.synth_query(6, 400)]
    offs = np.zeros(len(subs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    compare(engine, q, np.concatenate(subs), offs)


def test_repeats_long_subjects(engine):
    """Config 3 shape cut down: long subjects (median 800, up to 35,000), 20% homologs at 30-45%
    with low-complexity repeat blocks."""
    q, res, offs, _ = make(11, 1500, 300, len_median=800, len_max=35000, homolog_frac=0.2,
                           ident_min=0.30, ident_max=0.45, repeat_frac=0.05)
    compare(engine, q, res, offs)


def test_long_subject_rings(engine):
    """Subjects on both sides of the longest length the LDS diagonal ring takes (16,383 letters),
    and far beyond it, with copies of the query near their starts, middles and ends."""
    q = blz.synth_query(41, 700)
    subs = []
    for k, L in enumerate([16383, 16384, 16390, 40000, 65535]):
        s = blz.synth_query(500 + k, L).copy()
        for pos in (5, L // 2, L - len(q) - 1):
            s[pos:pos + len(q)] = q
        subs.append(s)
    subs.append(blz.synth_query(600, 300))
    offs = np.zeros(len(subs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    compare(engine, q, np.concatenate(subs), offs)


def test_subjects_past_16_bit_positions(engine):
    """Subjects of 65,536 letters and more (up to 150,000) are scanned in windows of word
    positions (a multiple of the diagonal ring's length, at most 32,768), with 16-bit offsets in
    the queue and lists and absolute positions in the ring.  Mutated query copies planted across
    the window boundaries, at the ends and every 9,000 letters give the oracle's lists, also when
    a tiny per-wave stage sends every long subject through the retry scan.  Seed variant 1 keeps
    16-bit positions and refuses such a database."""
    q = blz.synth_query(43, 700)
    rng = np.random.default_rng(5)
    subs = []
    for k, L in enumerate([65536, 70001, 150000]):
        s = blz.synth_query(700 + k, L).copy()
        for pos in list(range(5, L - 800, 9000)) + [31900, 32100, 32300, 64300, 64600, 96500, L - 701]:
            if pos + 700 > L:
                continue
            piece = q.copy()
            mut = rng.random(700) < 0.3
            piece[mut] = blz.synth_query(800 + pos, 700)[mut]
            s[pos:pos + 700] = piece
        subs.append(s)
    subs.append(blz.synth_query(601, 300))
    offs = np.zeros(len(subs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    res = np.concatenate(subs)
    saved = engine.get_tuning()
    if saved["seed_variant"] == 1:
        engine.upload_db(res, offs)
        qh = engine.prepare(q)
        try:
            with pytest.raises(blz.BlzError) as e:
                engine.run(qh)
            assert e.value.status == 1
        finally:
            engine.free_query(qh)
        return
    stats, nsub = compare(engine, q, res, offs)
    assert nsub == 3
    engine.set_tuning(**{**saved, "cap_stage": 8})
    try:
        compare(engine, q, res, offs)
    finally:
        engine.set_tuning(**saved)


def test_init_hit_cap(engine):
    """A subject with more init hits than the per-wave stage holds (6,704 > 4,096; low ungapped
    cutoff) and more HSPs than the wave pass keeps (1,799).  The batched kernel rescans it in its
    retry launch and the wide pass finishes it; the first seed kernel has no retry and the
    overflow queue finishes it: lists identical to the oracle's either way.  With overflow_mode 1
    the first kernel reports the pair instead (blz_overflow_drain) and every other list is exact."""
    q = blz.synth_query(77, 2000)
    subs = [blz.synth_query(78, 40000), blz.synth_query(79, 300), blz.synth_query(80, 5000)]
    offs = np.zeros(len(subs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    res = np.concatenate(subs)
    kw = dict(ungapped_cutoff=12, gapped_cutoff=22, x_gapped=12)
    stats, _ = compare(engine, q, res, offs, blz.default_params(**kw), oracle.params(**kw))
    assert stats["overflow_left"] == 0
    if engine.get_tuning()["seed_variant"] != 1:
        return
    assert stats["overflow_subjects"] == 1
    tune = engine.get_tuning()
    engine.set_tuning(overflow_mode=1)
    qh = engine.prepare(q, blz.default_params(**kw))
    try:
        got, _ = engine.run(qh, allow_overflow=True)
        assert engine.overflow_pairs() == [(0, 0)]
    finally:
        engine.free_query(qh)
        engine.set_tuning(**tune)
    counts, hsps, _ = oracle_run(q, res, offs, oracle.params(**kw))
    want = oracle.as_lists(counts, hsps)
    have = got.as_dict()
    assert 0 not in have and sorted(set(want) - {0}) == sorted(have)
    assert all(np.array_equal(want[k], have[k]) for k in have)


def test_empty_db(engine):
    q = blz.synth_query(1, 100)
    res = np.zeros(0, np.uint8)
    offs = np.zeros(1, np.uint64)
    eng = engine
    eng.upload_db(res, offs)
    qh = eng.prepare(q)
    got, stats = eng.run(qh)
    eng.free_query(qh)
    assert len(got.oids) == 0 and stats["word_hits"] == 0


def test_upload_rejects_bad_input(engine):
    """Letters outside NCBIstdaa (anywhere in the staged chunks) and subjects longer than
    16,777,215 letters are rejected with BLZ_EINVAL; nothing is left half-uploaded."""
    res = np.ones(1000, np.uint8)
    res[777] = 28
    offs = np.array([0, 500, 1000], np.uint64)
    with pytest.raises(blz.BlzError) as e:
        engine.upload_db(res, offs)
    assert e.value.status == 1
    big = np.ones(1 << 24, np.uint8)
    with pytest.raises(blz.BlzError) as e:
        engine.upload_db(big, np.array([0, 1 << 24], np.uint64))
    assert e.value.status == 1


def test_params_variants(engine):
    """Non-default cutoffs, X-drops and gap costs are honoured identically."""
    q, res, offs, _ = make(21, 400, 600, homolog_frac=0.1, ident_min=0.4, ident_max=0.6)
    kw = dict(window=30, x_ungapped=20, x_gapped=25, gap_open=9, gap_extend=2, ungapped_cutoff=35,
              gapped_cutoff=55)
    compare(engine, q, res, offs, blz.default_params(**kw), oracle.params(**kw))


@pytest.mark.parametrize("wave_classes", [1, 5])
def test_gapped_routing(engine, wave_classes):
    """Cost classes routed to the lane-per-subject DP (every class, or the single-hit classes
    while subjects with several init hits stay on waves) give the oracle's lists too."""
    saved = engine.get_tuning()
    q, res, offs, _ = make(31, 500, 20000, homolog_frac=0.05, ident_min=0.35, ident_max=0.7)
    engine.set_tuning(gapped_wave_classes=wave_classes)
    try:
        compare(engine, q, res, offs)
    finally:
        engine.set_tuning(**saved)


def test_config1_full_size(engine):
    """Config 1 at full size: 500 aa query vs 2,000,000 subjects, 1% homologs at 35-70%.
    The oracle runs multi-threaded on the host."""
    q, res, offs, hom = make(0, 500, 2_000_000, homolog_frac=0.01, ident_min=0.35, ident_max=0.70)
    stats, nsub = compare(engine, q, res, offs)
    assert stats["overflow_subjects"] == 0
    assert nsub >= 0.9 * (hom >= 0).sum()


def test_batch_matches_single_queries(engine):
    """Several queries of mixed length (tables in LDS and in global memory) in one batched run
    give, per query, exactly the oracle's lists."""
    if engine.get_tuning()["seed_variant"] == 1:
        pytest.skip("variant 1 runs one query at a time")
    qs = [blz.synth_query(100 + i, L) for i, L in enumerate([40, 300, 777, 1200, 3000])]
    spec = blz.synth_spec(seed=5, n_subjects=1500, homolog_frac=0.05, ident_min=0.35, ident_max=0.7)
    res, offs, _ = blz.synth_db(spec, qs)
    engine.upload_db(res, offs)
    qhs = [engine.prepare(q) for q in qs]
    try:
        outs, stats = engine.run_batch(qhs)
    finally:
        for h in qhs:
            engine.free_query(h)
    assert stats["seed_launches"] >= 1
    for q, got in zip(qs, outs):
        counts, hsps, _ = oracle.prelim_db(q, res, offs, nthreads=NTHREADS)
        want = oracle.as_lists(counts, hsps)
        have = got.as_dict()
        assert sorted(want) == sorted(have)
        for oid in want:
            assert np.array_equal(want[oid], have[oid]), (len(q), oid)


def test_packed_batch(engine):
    """Many short queries share seed scans (merged lookup tables over their concatenation);
    each query's lists and the summed counters equal the oracle's, with packing on or off, and a
    query with other parameters goes to a pack of its own."""
    if engine.get_tuning()["seed_variant"] == 1:
        pytest.skip("variant 1 runs one query at a time")
    lens = [3, 5, 11, 30, 47, 64, 90, 120, 150, 200, 230, 260, 300, 350, 380, 420, 40, 75, 9, 180]
    qs = [blz.synth_query(300 + i, L) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=31, n_subjects=1200, homolog_frac=0.08, ident_min=0.35, ident_max=0.8)
    res, offs, _ = blz.synth_db(spec, qs)
    engine.upload_db(res, offs)
    other = blz.default_params(word_threshold=12)
    p_other = oracle.params(word_threshold=12)
    params = [other if i == 7 else None for i in range(len(qs))]
    qhs = [engine.prepare(q, p) for q, p in zip(qs, params)]
    tune = engine.get_tuning()
    try:
        runs = {}
        for packing in (0, 1):
            engine.set_tuning(**{**tune, "query_packing": packing})
            runs[packing] = engine.run_batch(qhs)
            print(f"[case] packing {packing}: {runs[packing][1]['seed_launches']} seed launches", flush=True)
    finally:
        engine.set_tuning(**tune)
        for h in qhs:
            engine.free_query(h)
    totals = {k: 0 for k in ("word_hits", "ungapped_ext", "init_hits", "gapped_ext", "dp_cells")}
    for i, q in enumerate(qs):
        counts, hsps, ost = oracle.prelim_db(q, res, offs, p_other if i == 7 else None, nthreads=NTHREADS)
        for k in totals:
            totals[k] += ost[k]
        want = oracle.as_lists(counts, hsps)
        for packing, (outs, _) in runs.items():
            have = outs[i].as_dict()
            assert sorted(want) == sorted(have), (packing, len(q))
            for oid in want:
                assert np.array_equal(want[oid], have[oid]), (packing, len(q), oid)
    for packing, (_, stats) in runs.items():
        for k, v in totals.items():
            assert stats[k] == v, (packing, k)


def test_graph_replay_and_sinks(engine):
    """A batch captured as a hipGraph and replayed gives the eager lists on every replay, with
    another context's eager runs and host synchronisations between the replays; caller-sized
    sinks receive the same lists, and a too-small sink reports BLZ_ECAPACITY with the sizes."""
    if engine.get_tuning()["seed_variant"] == 1:
        pytest.skip("variant 1 runs one query at a time")
    qs = [blz.synth_query(700 + i, L) for i, L in enumerate([60, 200, 450, 900, 1300])]
    spec = blz.synth_spec(seed=17, n_subjects=1500, homolog_frac=0.1, ident_min=0.35, ident_max=0.8)
    res, offs, _ = blz.synth_db(spec, qs)
    engine.upload_db(res, offs)
    qhs = [engine.prepare(q) for q in qs]
    tune = engine.get_tuning()
    try:
        engine.set_tuning(**{**tune, "graph": 1})
        eager, st_e = engine.run_batch(qhs)
        engine.set_tuning(**{**tune, "graph": 0})
        with blz.Engine(0) as other:
            other.upload_db(res, offs)
            oq = other.prepare(qs[2])
            want_other, _ = other.run(oq)
            for rep in range(4):
                got, st = engine.run_batch(qhs)
                print(f"[case] replay {rep}: {st['seed_launches']} seed launches", flush=True)
                for k in ("word_hits", "ungapped_ext", "init_hits", "gapped_ext", "dp_cells", "hsps"):
                    assert st[k] == st_e[k], (rep, k)
                for a, b in zip(eager, got):
                    assert a.as_dict().keys() == b.as_dict().keys()
                    assert np.array_equal(a.hsps, b.hsps) and np.array_equal(a.oids, b.oids), rep
                o, _ = other.run(oq)  # eager work on another stream between replays
                assert np.array_equal(o.hsps, want_other.hsps)
            other.free_query(oq)
        # caller-sized sinks through the C ABI
        import ctypes as C
        lib = blz.load_library()
        n = len(qhs)
        keeps, sinks = [], (blz.Sink * n)()
        for i, r in enumerate(eager):
            sk, kp = engine._sink(len(r.oids), len(r.hsps))
            sinks[i] = sk
            keeps.append(kp)
        arr = (C.c_void_p * n)(*[q.value if isinstance(q, C.c_void_p) else q for q in qhs])
        assert lib.blz_prelim_run_batch(engine.ctx, arr, n, engine.db, sinks, None) == 0
        for i, r in enumerate(eager):
            oids, counts, hsps = keeps[i]
            assert sinks[i].n_subjects == len(r.oids) and sinks[i].n_hsps == len(r.hsps)
            assert np.array_equal(oids[: len(r.oids)], r.oids) and np.array_equal(hsps[: len(r.hsps)], r.hsps)
        big = max(range(n), key=lambda i: len(eager[i].hsps))
        small, _ = engine._sink(len(eager[big].oids), max(0, len(eager[big].hsps) - 1))
        assert lib.blz_result_copy(engine.ctx, big, C.byref(small)) == 4  # BLZ_ECAPACITY
        assert small.n_hsps == len(eager[big].hsps)
    finally:
        engine.set_tuning(**tune)
        for h in qhs:
            engine.free_query(h)
