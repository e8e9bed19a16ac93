"""C ABI and host logic that need no GPU: the library loads, exports every function the header
declares, refuses to run without a gfx950 device, and the synthetic generator is deterministic."""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

import blz

ROOT = pathlib.Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "blz_prelim.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(blz_[a-z0-9_]+)\s*\(", text)))


def test_exports_match_header():
    lib = blz.load_library()
    declared = header_functions()
    assert declared == sorted(blz.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def header_struct_fields(name):
    """Field names of a typedef struct in the header, in declaration order."""
    text = (ROOT / "include" / "blz_prelim.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct\s*\{([^}]*)\}\s*" + name + r"\s*;", text).group(1)
    names = []
    for decl in body.split(";"):
        if decl.strip():  # "type a, b" declares a and b
            names += [part.split()[-1].lstrip("*") for part in decl.split(",")]
    return names


@pytest.mark.parametrize("cname,pytype", [("blz_params", "Params"), ("blz_tuning", "Tuning"),
                                          ("blz_run_stats", "RunStats"), ("blz_hsp_sink", "Sink"),
                                          ("blz_server_config", "ServerConfig"),
                                          ("blz_request_info", "RequestInfo"),
                                          ("blz_server_stats", "ServerStats"), ("blz_karlin", "Karlin"),
                                          ("blz_stats_options", "StatsOptions"),
                                          ("blz_search_stats", "SearchStats")])
def test_binding_structs_match_header(cname, pytype):
    """The ctypes mirrors declare the header's fields, in the header's order (lambda_ stands for
    the C field lambda, a Python keyword)."""
    assert header_struct_fields(cname) == [f.rstrip("_") for f, _ in getattr(blz, pytype)._fields_]


def test_no_device_is_an_error_not_a_fallback():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(blz.BlzError) as e:
        blz.Engine(0)
    assert e.value.status == 3  # BLZ_EDEVICE


def test_default_params():
    p = blz.default_params()
    assert (p.word_threshold, p.window, p.x_ungapped, p.x_gapped, p.gap_open, p.gap_extend,
            p.ungapped_cutoff, p.gapped_cutoff) == (11, 40, 16, 38, 11, 1, 41, 67)


def test_synth_deterministic_and_sorted():
    q = blz.synth_query(0, 300)
    assert np.array_equal(q, blz.synth_query(0, 300))
    assert not np.array_equal(q, blz.synth_query(1, 300))
    assert set(np.unique(q)) <= {1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22}
    spec = blz.synth_spec(seed=0, n_subjects=5000)
    r1, o1, h1 = blz.synth_db(spec, [q])
    r2, o2, h2 = blz.synth_db(spec, [q])
    assert np.array_equal(r1, r2) and np.array_equal(o1, o2) and np.array_equal(h1, h2)
    lens = np.diff(o1.astype(np.int64))
    assert (np.diff(lens) >= 0).all() and lens.min() >= 30 and lens.max() <= 5000
    assert 200 <= np.median(lens) <= 400
    assert 0.01 < (h1 >= 0).mean() < 0.03


def test_synth_homolog_identity():
    """Planted homologs without indels carry the requested identity to the query segment."""
    q = blz.synth_query(4, 400)
    spec = blz.synth_spec(seed=4, n_subjects=400, homolog_frac=1.0, ident_min=0.5, ident_max=0.5,
                          indel_rate=0.0, len_min=400, len_max=400)
    res, offs, hom = blz.synth_db(spec, [q])
    ids = [(res[int(offs[i]):int(offs[i + 1])] == q).mean() for i in range(len(offs) - 1)]
    assert abs(np.mean(ids) - 0.5) < 0.03


def test_synth_config_sizes():
    """Residue totals of the benchmark configurations (log-normal mean = median * e^(s^2/2))."""
    lib = blz.load_library()
    n, r = C.c_int64(), C.c_uint64()
    spec = blz.synth_spec(seed=0, n_subjects=200_000)
    assert lib.blz_synth_db_size(C.byref(spec), C.byref(n), C.byref(r)) == 0
    mean = r.value / n.value
    assert 400 < mean < 425  # 300 * exp(0.32) = 413 before clipping


def test_bad_inputs_rejected_without_device():
    lib = blz.load_library()
    h = C.c_void_p()
    assert lib.blz_ctx_create(-1, C.byref(h)) == 3
    assert lib.blz_db_upload(None, None, None, 0, C.byref(h)) == 1
    assert lib.blz_query_prepare(None, None, 0, None, C.byref(h)) == 1
    assert lib.blz_prelim_run(None, None, None, None, None) == 1
    assert lib.blz_prelim_run_noalloc(None, None, None, None, None) == 1


def test_traceback_result_views():
    """TbResult: pairs in call order, each pair's HSPs with op_first rebased to its own runs."""
    pa = np.zeros(3, blz.TB_PAIR_DTYPE)
    pa["query"], pa["oid"], pa["n_hsps"] = [0, 0, 1], [5, 9, 5], [1, 0, 2]
    tb = np.zeros(3, blz.TB_DTYPE)
    tb["op_first"], tb["op_count"], tb["score"] = [0, 2, 3], [2, 1, 2], [10, 20, 15]
    r = blz.TbResult(2, pa, np.array([1, 0, 2], np.int32), tb, np.arange(5, dtype=np.uint32))
    t, runs = r[0][5]
    assert t["score"].tolist() == [10] and runs.tolist() == [0, 1]
    assert len(r[0][9][0]) == 0 and len(r[0][9][1]) == 0
    t, runs = r[1][5]
    assert t["op_first"].tolist() == [0, 1] and runs.tolist() == [2, 3, 4]
    assert sorted(r[0].keys()) == [5, 9] and 5 in r[1] and len(r) == 2
