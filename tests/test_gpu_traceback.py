"""Traceback stage on the GPU (blz_traceback) against the oracle's (oracle/traceback.c).

Every traced HSP -- score, extent, gapped start, identities, columns, gaps, gap openings -- and
its edit script must equal the oracle's, pair for pair, for the hit lists of real prelim runs:
the config-0 database with one 300-letter query, and the reference's 40 benchmark queries (plus
three with ambiguity letters) against a database with ambiguity letters planted.  The wide pass
and the edit-script rerun are forced by their capacity limits and must not change a bit.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import blz
import oracle

pytestmark = pytest.mark.gpu

NTHREADS = len(os.sched_getaffinity(0))
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "queries_ncbistdaa.txt")
AMBIGUOUS = np.array([2, 21, 23, 24, 25, 26, 27], np.uint8)  # B X Z U * O J


def log(msg):
    print(f"[case] {msg}", flush=True)


def golden_queries():
    out = []
    for line in open(GOLDEN):
        _, *codes = line.split()
        out.append(np.array(list(map(int, codes)), np.uint8))
    return out


def oracle_pairs(queries, res, offs, lists, x_final, cutoff):
    """{(query index, oid): (traced HSPs, runs)} from the oracle, on the host's threads."""
    jobs = [(qi, oid, hs) for qi, r in enumerate(lists) for oid, hs in r.as_dict().items()]

    def one(job):
        qi, oid, hs = job
        s = res[int(offs[oid]):int(offs[oid + 1])]
        return (qi, oid), oracle.traceback_subject(queries[qi], s, hs, x_final=x_final, cutoff=cutoff)

    with ThreadPoolExecutor(NTHREADS) as ex:
        return dict(ex.map(one, jobs))


def compare(want, got):
    n = 0
    for (qi, oid), (wt, wr) in want.items():
        gt, gr = got[qi][oid]
        assert gt.tolist() == wt.tolist(), (qi, oid, gt, wt)
        assert gr.tolist() == wr.tolist(), (qi, oid)
        n += len(wt)
    return n


def prelim_and_hitlists(eng, queries, res, offs, expect=10.0):
    hs = [eng.prepare(q) for q in queries]
    outs, _ = eng.run_batch(hs)
    lists, stats = [], []
    for i, q in enumerate(queries):
        ss = blz.search_setup(q, int(offs[-1]), len(offs) - 1, blz.stats_options(expect=expect))[1]
        lst, _ = eng.hitlist(i, ss)
        lists.append(lst)
        stats.append(ss)
    return hs, lists, stats


def test_self_alignment_known_answer():
    """A subject equal to the query: one HSP over the whole sequence, every column an identity,
    one run of aligned pairs."""
    q = blz.synth_query(5, 240)
    spec = blz.synth_spec(seed=5, n_subjects=200, homolog_frac=0.0)
    res, offs, _ = blz.synth_db(spec, [q])
    res = np.concatenate([res, q])
    offs = np.concatenate([offs, [offs[-1] + len(q)]]).astype(np.uint64)
    oid = len(offs) - 2
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        h = eng.prepare(q)
        r, _ = eng.run(h)
        lst = blz.Result(np.array([oid]), np.array([r.as_dict()[oid].shape[0]], np.int32), r.as_dict()[oid])
        got, st = eng.traceback([h], [lst])
        eng.free_query(h)
    t, runs = got[0][oid]
    assert len(t) == 1 and (t[0]["q_off"], t[0]["q_end"], t[0]["s_off"], t[0]["s_end"]) == (0, 240, 0, 240)
    assert (t[0]["num_ident"], t[0]["align_len"], t[0]["gaps"], t[0]["gap_opens"]) == (240, 240, 0, 0)
    assert runs.tolist() == [(blz.OP_SUB << 28) | 240]
    assert st["hsps"] == 1 and st["alignments"] == 2


def test_config0_hitlist_equals_oracle():
    q = blz.synth_query(0, 300)
    spec = blz.synth_spec(seed=0, n_subjects=10_000)
    res, offs, _ = blz.synth_db(spec, [q])
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        hs, lists, stats = prelim_and_hitlists(eng, [q], res, offs)
        x_final, cutoff = stats[0].x_final, stats[0].reap_score
        got, st = eng.traceback(hs, lists, blz.tb_params(x_dropoff=x_final), cutoffs=[cutoff])
        for h in hs:
            eng.free_query(h)
    log(f"config 0: {st['pairs']} pairs, {st['hsps']} traced HSPs, {st['dp_cells']} cells, "
        f"{st['ms_main']:.2f} ms main pass, {st['wide_pairs']} wide")
    want = oracle_pairs([q], res, offs, lists, x_final, cutoff)
    n = compare(want, got)
    assert n == st["hsps"] > 0
    for oid, (t, _) in got[0].items():
        for rec in t:
            assert blz.evalue(stats[0], int(rec["score"])) <= 10.0


def test_reference_queries_with_ambiguity_letters():
    rng = np.random.default_rng(19)
    queries = golden_queries()
    for k in (0, 7, 21):
        v = queries[k].copy()
        pos = rng.random(len(v)) < 0.02
        v[pos] = rng.choice(AMBIGUOUS, int(pos.sum()))
        queries.append(v)
    spec = blz.synth_spec(seed=1, n_subjects=10_000, homolog_frac=0.05)
    res, offs, _ = blz.synth_db(spec, queries)
    res = res.copy()
    pos = rng.random(res.size) < 0.01
    res[pos] = rng.choice(AMBIGUOUS, int(pos.sum()))
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        hs, lists, stats = prelim_and_hitlists(eng, queries, res, offs)
        got, st = eng.traceback(hs, lists, blz.tb_params(x_dropoff=stats[0].x_final))
        log(f"{len(queries)} queries: {st['pairs']} pairs, {st['hsps']} traced HSPs, {st['dp_cells']} cells, "
            f"{st['ms_main']:.2f} ms main + {st['ms_wide']:.2f} ms wide ({st['wide_pairs']} pairs)")
        # four pairs per wave (16-lane segments) instead of one
        eng.set_tuning(tb_lanes=16)
        wave, wst = eng.traceback(hs, lists, blz.tb_params(x_dropoff=stats[0].x_final))
        log(f"16 lanes per pair: {wst['ms_main']:.2f} ms main, {wst['wide_pairs']} wide pairs")
        # forced: all but the shortest pairs through the wide pass, and an edit-script buffer that
        # must grow
        eng.set_tuning(tb_lanes=0, cap_tb_bytes=4096, cap_tb_runs=64)
        forced, fst = eng.traceback(hs, lists, blz.tb_params(x_dropoff=stats[0].x_final))
        log(f"forced: {fst['wide_pairs']} wide pairs, {fst['reruns']} reruns")
        for h in hs:
            eng.free_query(h)
    assert fst["wide_pairs"] > 0.5 * st["pairs"] and fst["reruns"] >= 1
    for other in (wave, forced):
        for qi in range(len(queries)):
            assert other[qi].keys() == got[qi].keys()
            for oid in got[qi]:
                assert other[qi][oid][0].tolist() == got[qi][oid][0].tolist()
                assert other[qi][oid][1].tolist() == got[qi][oid][1].tolist()
    want = oracle_pairs(queries, res, offs, lists, stats[0].x_final, 0)
    n = compare(want, got)
    log(f"{len(want)} pairs, {n} traced HSPs equal the oracle")
    assert n == st["hsps"] > 0


def test_edge_pairs_and_bad_input():
    """Pairs without HSPs; an HSP contained in one traced before it; gapped starts outside the
    pair are refused."""
    q = blz.synth_query(7, 150)
    spec = blz.synth_spec(seed=7, n_subjects=500, homolog_frac=0.2)
    res, offs, _ = blz.synth_db(spec, [q])
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        h = eng.prepare(q)
        r, _ = eng.run(h)
        d = r.as_dict()
        oid = max(d, key=lambda k: d[k]["score"][0])
        top = d[oid][:1]
        s = res[int(offs[oid]):int(offs[oid + 1])]
        wt, wr = oracle.traceback_subject(q, s, top)
        assert wt["q_end"][0] - wt["q_off"][0] > 20 and wt["s_end"][0] - wt["s_off"][0] > 20
        inner = top.copy()  # inside the traced HSP's box, scoring lower: skipped
        inner["score"] = wt["score"][0] - 3
        inner["q_off"], inner["q_end"] = wt["q_off"][0] + 5, wt["q_end"][0] - 5
        inner["s_off"], inner["s_end"] = wt["s_off"][0] + 5, wt["s_end"][0] - 5
        inner["q_gs"], inner["s_gs"] = wt["q_off"][0] + 10, wt["s_off"][0] + 10
        both = np.concatenate([top, inner])
        lst = blz.Result(np.array([oid, 3, oid]), np.array([1, 0, 2], np.int32), np.concatenate([top, both]))
        got, st = eng.traceback([h], [lst])
        assert st["pairs"] == 3 and st["alignments"] == 4  # the inner HSP is not aligned again
        assert got[0][3][0].shape == (0,)
        bad = top.copy()
        bad["s_gs"] = int(offs[oid + 1] - offs[oid])
        with pytest.raises(blz.BlzError) as e:
            eng.traceback([h], [blz.Result(np.array([oid]), np.array([1], np.int32), bad)])
        assert e.value.status == 1
        with pytest.raises(blz.BlzError):
            eng.traceback([h], [blz.Result(np.array([len(offs) - 1]), np.array([0], np.int32), top[:0])])
        eng.free_query(h)
    assert got[0][oid][0].tolist() == wt.tolist() and got[0][oid][1].tolist() == wr.tolist()


def test_subjects_past_16_bit_positions():
    """Subjects of 65,536 letters and more (up to 150,000) with mutated query copies: every prelim
    list traced, HSPs and edit scripts equal to the oracle's."""
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
    offs = np.zeros(len(subs) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum([len(s) for s in subs])
    res = np.concatenate(subs)
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        h = eng.prepare(q)
        r, _ = eng.run(h)
        got, st = eng.traceback([h], [r])
        eng.free_query(h)
    n = 0
    for oid, hs in r.as_dict().items():
        wt, wr = oracle.traceback_subject(q, res[int(offs[oid]):int(offs[oid + 1])], hs)
        gt, gr = got[0][oid]
        assert gt.tolist() == wt.tolist() and gr.tolist() == wr.tolist(), oid
        n += len(wt)
    log(f"long subjects: {len(r.oids)} pairs, {n} traced HSPs equal the oracle")
    assert len(r.oids) == 3 and n > 10


@pytest.mark.parametrize("go,ge,x", [(10, 2, 40), (13, 1, 90), (7, 3, 30)])
def test_other_gap_costs_and_xdrops(go, ge, x):
    """Gap costs other than 11/1 (an extension of 2 or 3 takes the division and the multiples of
    the extension in the band rules) and other final X-drops, prelim lists from the same costs,
    every pair against the oracle."""
    q = blz.synth_query(60 + go, 350)
    spec = blz.synth_spec(seed=60 + go, n_subjects=4000, homolog_frac=0.08, ident_min=0.3, ident_max=0.8,
                          indel_rate=0.03)
    res, offs, _ = blz.synth_db(spec, [q])
    p = blz.default_params(gap_open=go, gap_extend=ge)
    with blz.Engine(0) as eng:
        eng.upload_db(res, offs)
        h = eng.prepare(q, p)
        r, _ = eng.run(h)
        got, st = eng.traceback([h], [r], blz.tb_params(x_dropoff=x, gap_open=go, gap_extend=ge))
        eng.free_query(h)
    n = 0
    for oid, hs in r.as_dict().items():
        wt, wr = oracle.traceback_subject(q, res[int(offs[oid]):int(offs[oid + 1])], hs, x_final=x, gap_open=go,
                                          gap_extend=ge)
        gt, gr = got[0][oid]
        assert gt.tolist() == wt.tolist() and gr.tolist() == wr.tolist(), oid
        n += len(wt)
    log(f"gap {go}/{ge}, X {x}: {len(r.oids)} pairs, {n} traced HSPs equal the oracle")
    assert n > 10
