"""Python binding of the MI355X BLASTP prelim engine (ctypes over include/blz_prelim.h).

Host-side mirror of the reference's hook for this path: the reference calls
``BLAZE_GPU(seq_src, threadIndex, totalThreads, queryLen, survivingSeqs, num_survivors,
ungapped_cutoff, gapped_cutoff)`` (BLAZE_gpu/gpu_cpu_common.h:43) and then re-runs NCBI's CPU
prelim core on the survivors (BLAZE_ncbi/blast_engine.c:2983-2996).  Here one call returns the
per-subject HSP lists that core would write to the HSP stream (blast_engine.c:2499).

The engine only runs on a gfx950 device through ``libblaze_amd.so``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libblaze_amd.so"

HSP_DTYPE = np.dtype(
    [("score", "<i4"), ("q_off", "<i4"), ("q_end", "<i4"), ("s_off", "<i4"), ("s_end", "<i4"),
     ("q_gs", "<i4"), ("s_gs", "<i4")]
)

# blz_tb_hsp: a traced HSP; op_first/op_count index its edit script (runs op << 28 | length)
TB_DTYPE = np.dtype(
    [("score", "<i4"), ("q_off", "<i4"), ("q_end", "<i4"), ("s_off", "<i4"), ("s_end", "<i4"),
     ("q_gs", "<i4"), ("s_gs", "<i4"), ("num_ident", "<i4"), ("align_len", "<i4"), ("gaps", "<i4"),
     ("gap_opens", "<i4"), ("op_first", "<i4"), ("op_count", "<i4")]
)
OP_SUB, OP_GAP_Q, OP_GAP_S = 0, 1, 2  # aligned pair; subject letter vs gap; query letter vs gap
TB_PAIR_DTYPE = np.dtype([("query", "<i4"), ("n_hsps", "<i4"), ("oid", "<i8")])

STATUS = {0: "BLZ_OK", 1: "BLZ_EINVAL", 2: "BLZ_ENOMEM", 3: "BLZ_EDEVICE", 4: "BLZ_ECAPACITY",
          5: "BLZ_EOVERFLOW", 6: "BLZ_ETIMEOUT"}


class BlzError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [("word_threshold", C.c_int32), ("window", C.c_int32), ("x_ungapped", C.c_int32),
                ("x_gapped", C.c_int32), ("gap_open", C.c_int32), ("gap_extend", C.c_int32),
                ("ungapped_cutoff", C.c_int32), ("gapped_cutoff", C.c_int32), ("reap_score", C.c_int32)]


class Karlin(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("K", C.c_double), ("logK", C.c_double), ("H", C.c_double)]


class StatsOptions(C.Structure):
    _fields_ = [("expect", C.c_double), ("prelim_evalue", C.c_double), ("x_ungapped_bits", C.c_double),
                ("x_gapped_bits", C.c_double), ("x_final_bits", C.c_double), ("gap_trigger_bits", C.c_double),
                ("gap_open", C.c_int32), ("gap_extend", C.c_int32), ("word_threshold", C.c_int32),
                ("window", C.c_int32), ("query_composition", C.c_int32), ("max_target_seqs", C.c_int32),
                ("composition_based_stats", C.c_int32), ("reserved", C.c_int32)]


class SearchStats(C.Structure):
    _fields_ = [("ungapped", Karlin), ("gapped", Karlin), ("alpha", C.c_double), ("beta", C.c_double),
                ("length_adjustment", C.c_int32), ("gap_trigger", C.c_int32), ("eff_query_len", C.c_int64),
                ("eff_db_len", C.c_int64), ("eff_searchsp", C.c_double), ("prelim_evalue", C.c_double),
                ("reap_score", C.c_int32), ("x_final", C.c_int32), ("hitlist_size", C.c_int32),
                ("reserved", C.c_int32)]


class RunStats(C.Structure):
    _fields_ = [("subjects", C.c_uint64), ("residues", C.c_uint64), ("word_hits", C.c_uint64),
                ("ungapped_ext", C.c_uint64), ("init_hits", C.c_uint64),
                ("gapped_subjects", C.c_uint64), ("gapped_ext", C.c_uint64),
                ("dp_cells", C.c_uint64), ("hsps", C.c_uint64), ("overflow_subjects", C.c_uint64),
                ("ms_seed", C.c_double), ("ms_gapped", C.c_double), ("ms_total", C.c_double),
                ("seed_launches", C.c_int32), ("gapped_launches", C.c_int32),
                ("overflow_left", C.c_uint64), ("reruns", C.c_int32), ("drain_rounds", C.c_int32),
                ("ms_drain", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("seed_variant", C.c_int32), ("seed_waves", C.c_int32), ("seed_blocks_per_cu", C.c_int32),
                ("gapped_blocks", C.c_int32), ("gapped_threads", C.c_int32), ("seed_table", C.c_int32),
                ("gapped_wave_classes", C.c_int32), ("query_packing", C.c_int32), ("pack_len", C.c_int32),
                ("bin_ring", C.c_int32), ("graph", C.c_int32), ("gapped_lanes", C.c_int32),
                ("subject_groups", C.c_int32), ("overflow_mode", C.c_int32), ("cap_stage", C.c_int32),
                ("cap_retry", C.c_int32), ("cap_ovf", C.c_int32), ("cap_wide_ring", C.c_int32),
                ("cap_wide_saved", C.c_int32), ("cap_work", C.c_int64), ("cap_pool", C.c_int64),
                ("cap_hsps", C.c_int64), ("cap_tb_bytes", C.c_int64), ("cap_tb_runs", C.c_int64),
                ("tb_lanes", C.c_int32), ("gapped_width", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_batch", C.c_int32), ("max_batch_letters", C.c_int64),
                ("linger_us", C.c_int32), ("max_batch_work", C.c_int64), ("overlap_ms", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("batch_queries", C.c_int32), ("latency_us", C.c_double),
                ("queue_us", C.c_double), ("n_subjects", C.c_int64), ("n_hsps", C.c_int64),
                ("n_overflow", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("residues", C.c_uint64), ("device_ms", C.c_double), ("reruns", C.c_uint64),
                ("work_rate", C.c_double), ("pair_rate", C.c_double)]


class TbParams(C.Structure):
    _fields_ = [("x_dropoff", C.c_int32), ("gap_open", C.c_int32), ("gap_extend", C.c_int32),
                ("cutoff", C.c_int32)]


class TbStats(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("prelim_hsps", C.c_uint64), ("hsps", C.c_uint64),
                ("alignments", C.c_uint64), ("dp_cells", C.c_uint64), ("runs", C.c_uint64),
                ("wide_pairs", C.c_uint64), ("reruns", C.c_int32), ("reserved", C.c_int32),
                ("ms_main", C.c_double), ("ms_wide", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Sink(C.Structure):
    _fields_ = [("oids", C.POINTER(C.c_int64)), ("counts", C.POINTER(C.c_int32)),
                ("subject_cap", C.c_int64), ("hsps", C.c_void_p), ("hsp_cap", C.c_int64),
                ("n_subjects", C.c_int64), ("n_hsps", C.c_int64)]


class SynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_subjects", C.c_int64), ("len_median", C.c_double),
                ("len_sigma", C.c_double), ("len_min", C.c_int32), ("len_max", C.c_int32),
                ("homolog_frac", C.c_double), ("ident_min", C.c_double), ("ident_max", C.c_double),
                ("indel_rate", C.c_double), ("indel_mean", C.c_double), ("repeat_frac", C.c_double),
                ("repeat_period_min", C.c_int32), ("repeat_period_max", C.c_int32),
                ("sort_by_length", C.c_int32)]


# every symbol include/blz_prelim.h declares
EXPORTS = [
    "blz_ctx_create", "blz_ctx_destroy", "blz_last_error", "blz_ctx_stream", "blz_default_params",
    "blz_db_upload", "blz_db_free", "blz_db_num_subjects", "blz_db_num_residues",
    "blz_query_prepare", "blz_query_free", "blz_prelim_run", "blz_prelim_run_batch",
    "blz_synth_default", "blz_synth_query", "blz_synth_db_size", "blz_synth_db_fill",
    "blz_ctx_set_tuning", "blz_ctx_get_tuning", "blz_result_sizes", "blz_result_copy", "blz_overflow_drain",
    "blz_server_default_config", "blz_server_create", "blz_server_destroy", "blz_server_submit",
    "blz_server_wait", "blz_server_copy", "blz_server_release", "blz_server_get_stats",
    "blz_stats_default_options", "blz_karlin_ungapped", "blz_karlin_gapped", "blz_search_setup", "blz_evalue",
    "blz_result_hitlist", "blz_prelim_run_noalloc", "blz_server_overflow",
    "blz_blastdb_open", "blz_blastdb_close", "blz_blastdb_last_error", "blz_blastdb_num_seqs",
    "blz_blastdb_num_residues", "blz_blastdb_max_length", "blz_blastdb_title", "blz_blastdb_sequence",
    "blz_blastdb_header", "blz_blastdb_gather", "blz_db_upload_blastdb",
    "blz_tb_default_params", "blz_traceback", "blz_traceback_sizes", "blz_traceback_copy",
]

_lib = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libblaze_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(p))
    vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
    lib.blz_ctx_create.argtypes = [i32, C.POINTER(vp)]
    lib.blz_ctx_create.restype = C.c_int
    lib.blz_ctx_destroy.argtypes = [vp]
    lib.blz_last_error.argtypes = [vp]
    lib.blz_last_error.restype = C.c_char_p
    lib.blz_ctx_stream.argtypes = [vp]
    lib.blz_ctx_stream.restype = vp
    lib.blz_default_params.argtypes = [C.POINTER(Params)]
    lib.blz_db_upload.argtypes = [vp, vp, vp, i64, C.POINTER(vp)]
    lib.blz_db_upload.restype = C.c_int
    lib.blz_db_free.argtypes = [vp]
    lib.blz_db_num_subjects.argtypes = [vp]
    lib.blz_db_num_subjects.restype = i64
    lib.blz_db_num_residues.argtypes = [vp]
    lib.blz_db_num_residues.restype = u64
    lib.blz_query_prepare.argtypes = [vp, vp, i32, C.POINTER(Params), C.POINTER(vp)]
    lib.blz_query_prepare.restype = C.c_int
    lib.blz_query_free.argtypes = [vp]
    lib.blz_prelim_run.argtypes = [vp, vp, vp, C.POINTER(Sink), C.POINTER(RunStats)]
    lib.blz_prelim_run.restype = C.c_int
    lib.blz_prelim_run_batch.argtypes = [vp, C.POINTER(vp), i32, vp, C.POINTER(Sink), C.POINTER(RunStats)]
    lib.blz_prelim_run_batch.restype = C.c_int
    lib.blz_ctx_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.blz_ctx_set_tuning.restype = C.c_int
    lib.blz_ctx_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.blz_ctx_get_tuning.restype = C.c_int
    lib.blz_result_sizes.argtypes = [vp, i32, C.POINTER(i64), C.POINTER(i64)]
    lib.blz_result_sizes.restype = C.c_int
    lib.blz_result_copy.argtypes = [vp, i32, C.POINTER(Sink)]
    lib.blz_result_copy.restype = C.c_int
    lib.blz_overflow_drain.argtypes = [vp, vp, i64, C.POINTER(i64)]
    lib.blz_overflow_drain.restype = C.c_int
    lib.blz_server_default_config.argtypes = [C.POINTER(ServerConfig)]
    lib.blz_server_create.argtypes = [i32, vp, C.POINTER(ServerConfig), C.POINTER(Tuning), C.POINTER(vp)]
    lib.blz_server_create.restype = C.c_int
    lib.blz_server_destroy.argtypes = [vp]
    lib.blz_server_submit.argtypes = [vp, vp, i32, C.POINTER(Params), C.POINTER(u64)]
    lib.blz_server_submit.restype = C.c_int
    lib.blz_server_wait.argtypes = [vp, u64, i32, C.POINTER(RequestInfo)]
    lib.blz_server_wait.restype = C.c_int
    lib.blz_server_copy.argtypes = [vp, u64, C.POINTER(Sink)]
    lib.blz_server_copy.restype = C.c_int
    lib.blz_server_overflow.argtypes = [vp, u64, vp, i64, C.POINTER(i64)]
    lib.blz_server_overflow.restype = C.c_int
    lib.blz_server_release.argtypes = [vp, u64]
    lib.blz_server_release.restype = C.c_int
    lib.blz_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
    lib.blz_server_get_stats.restype = C.c_int
    lib.blz_synth_default.argtypes = [C.POINTER(SynthSpec)]
    lib.blz_synth_query.argtypes = [u64, i32, vp]
    lib.blz_synth_db_size.argtypes = [C.POINTER(SynthSpec), C.POINTER(i64), C.POINTER(u64)]
    lib.blz_synth_db_size.restype = C.c_int
    lib.blz_synth_db_fill.argtypes = [C.POINTER(SynthSpec), C.POINTER(vp), vp, i32, vp, vp, vp]
    lib.blz_synth_db_fill.restype = C.c_int
    lib.blz_stats_default_options.argtypes = [C.POINTER(StatsOptions)]
    lib.blz_karlin_ungapped.argtypes = [vp, C.POINTER(Karlin)]
    lib.blz_karlin_ungapped.restype = C.c_int
    lib.blz_karlin_gapped.argtypes = [i32, i32, C.POINTER(Karlin), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.blz_karlin_gapped.restype = C.c_int
    lib.blz_search_setup.argtypes = [C.POINTER(StatsOptions), vp, i32, u64, i64, C.POINTER(Params),
                                     C.POINTER(SearchStats)]
    lib.blz_search_setup.restype = C.c_int
    lib.blz_evalue.argtypes = [C.POINTER(SearchStats), i32]
    lib.blz_evalue.restype = C.c_double
    lib.blz_result_hitlist.argtypes = [vp, i32, C.POINTER(SearchStats), C.POINTER(Sink), vp]
    lib.blz_result_hitlist.restype = C.c_int
    lib.blz_prelim_run_noalloc.argtypes = [vp, vp, vp, C.POINTER(Sink), C.POINTER(RunStats)]
    lib.blz_prelim_run_noalloc.restype = C.c_int
    lib.blz_blastdb_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.blz_blastdb_open.restype = C.c_int
    lib.blz_blastdb_close.argtypes = [vp]
    lib.blz_blastdb_last_error.restype = C.c_char_p
    lib.blz_blastdb_num_seqs.argtypes = [vp]
    lib.blz_blastdb_num_seqs.restype = i64
    lib.blz_blastdb_num_residues.argtypes = [vp]
    lib.blz_blastdb_num_residues.restype = u64
    lib.blz_blastdb_max_length.argtypes = [vp]
    lib.blz_blastdb_max_length.restype = i32
    lib.blz_blastdb_title.argtypes = [vp]
    lib.blz_blastdb_title.restype = C.c_char_p
    lib.blz_blastdb_sequence.argtypes = [vp, i64, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(i32)]
    lib.blz_blastdb_sequence.restype = C.c_int
    lib.blz_blastdb_header.argtypes = [vp, i64, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(i64)]
    lib.blz_blastdb_header.restype = C.c_int
    lib.blz_blastdb_gather.argtypes = [vp, i32, vp, vp, vp]
    lib.blz_blastdb_gather.restype = C.c_int
    lib.blz_db_upload_blastdb.argtypes = [vp, vp, i32, C.POINTER(vp), vp]
    lib.blz_db_upload_blastdb.restype = C.c_int
    lib.blz_tb_default_params.argtypes = [C.POINTER(TbParams)]
    lib.blz_traceback.argtypes = [vp, vp, C.POINTER(vp), i32, vp, vp, i64, vp, C.POINTER(TbParams),
                                  C.POINTER(TbStats)]
    lib.blz_traceback.restype = C.c_int
    lib.blz_traceback_sizes.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
    lib.blz_traceback_sizes.restype = C.c_int
    lib.blz_traceback_copy.argtypes = [vp, vp, i64, vp, i64, vp, i64]
    lib.blz_traceback_copy.restype = C.c_int
    if path is None:
        _lib = lib
    return lib


class BlastDB:
    """A BLAST protein database (makeblastdb format 4 or 5) opened with blz_blastdb_open."""

    def __init__(self, path):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.blz_blastdb_open(str(path).encode(), C.byref(h))
        if st:
            raise BlzError(st, f"blz_blastdb_open: {self.lib.blz_blastdb_last_error().decode()}")
        self.h = h

    def close(self):
        if self.h:
            self.lib.blz_blastdb_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def num_seqs(self) -> int:
        return self.lib.blz_blastdb_num_seqs(self.h)

    @property
    def num_residues(self) -> int:
        return self.lib.blz_blastdb_num_residues(self.h)

    @property
    def max_length(self) -> int:
        return self.lib.blz_blastdb_max_length(self.h)

    @property
    def title(self) -> str:
        return self.lib.blz_blastdb_title(self.h).decode()

    def sequence(self, oid: int) -> np.ndarray:
        p, n = C.POINTER(C.c_uint8)(), C.c_int32()
        st = self.lib.blz_blastdb_sequence(self.h, oid, C.byref(p), C.byref(n))
        if st:
            raise BlzError(st, "blz_blastdb_sequence")
        return np.ctypeslib.as_array(p, shape=(max(1, n.value),))[: n.value].copy()

    def header(self, oid: int) -> bytes:
        p, n = C.POINTER(C.c_uint8)(), C.c_int64()
        st = self.lib.blz_blastdb_header(self.h, oid, C.byref(p), C.byref(n))
        if st:
            raise BlzError(st, "blz_blastdb_header")
        return bytes(np.ctypeslib.as_array(p, shape=(max(1, n.value),))[: n.value]) if n.value else b""

    def gather(self, sort_by_length: bool = False):
        """(residues, offsets, perm): the sequences in OID order or ascending length."""
        n = self.num_seqs
        res = np.empty(self.num_residues + 1, np.uint8)
        offs = np.empty(n + 1, np.uint64)
        perm = np.empty(max(1, n), np.int64)
        st = self.lib.blz_blastdb_gather(self.h, int(sort_by_length), res.ctypes.data, offs.ctypes.data,
                                         perm.ctypes.data)
        if st:
            raise BlzError(st, f"blz_blastdb_gather: {self.lib.blz_blastdb_last_error().decode()}")
        return res[: self.num_residues], offs, perm[:n]


# ---------------------------------------------------------------------------------------
# search statistics
# ---------------------------------------------------------------------------------------
def stats_options(**kw) -> StatsOptions:
    o = StatsOptions()
    load_library().blz_stats_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def karlin_ungapped(freq=None) -> Karlin:
    k = Karlin()
    f = None if freq is None else np.ascontiguousarray(freq, dtype=np.float64)
    st = load_library().blz_karlin_ungapped(None if f is None else f.ctypes.data, C.byref(k))
    if st:
        raise BlzError(st, "blz_karlin_ungapped")
    return k


def karlin_gapped(gap_open: int, gap_extend: int):
    k, a, b = Karlin(), C.c_double(), C.c_double()
    st = load_library().blz_karlin_gapped(gap_open, gap_extend, C.byref(k), C.byref(a), C.byref(b))
    if st:
        raise BlzError(st, "blz_karlin_gapped")
    return k, a.value, b.value


def search_setup(query, db_residues: int, db_seqs: int, options: StatsOptions | None = None):
    """(Params, SearchStats) of one query against a database of that size."""
    q = np.ascontiguousarray(query, dtype=np.uint8)
    o = options or stats_options()
    p, s = Params(), SearchStats()
    st = load_library().blz_search_setup(C.byref(o), q.ctypes.data, len(q), db_residues, db_seqs, C.byref(p),
                                         C.byref(s))
    if st:
        raise BlzError(st, "blz_search_setup")
    return p, s


def evalue(stats: SearchStats, score: int) -> float:
    return load_library().blz_evalue(C.byref(stats), int(score))


def default_params(**kw) -> Params:
    p = Params()
    load_library().blz_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------------------------------
# synthetic data
# ---------------------------------------------------------------------------------------
def synth_spec(**kw) -> SynthSpec:
    s = SynthSpec()
    load_library().blz_synth_default(C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def synth_query(seed: int, length: int) -> np.ndarray:
    out = np.empty(length, dtype=np.uint8)
    load_library().blz_synth_query(seed, length, out.ctypes.data)
    return out


def synth_db(spec: SynthSpec, queries: list[np.ndarray]):
    """Returns (residues uint8, offsets uint64[n+1], homolog_of int32[n])."""
    lib = load_library()
    n = C.c_int64()
    r = C.c_uint64()
    st = lib.blz_synth_db_size(C.byref(spec), C.byref(n), C.byref(r))
    if st:
        raise BlzError(st, "synth_db_size")
    residues = np.empty(r.value + 64, dtype=np.uint8)
    offsets = np.empty(n.value + 1, dtype=np.uint64)
    hom = np.empty(n.value, dtype=np.int32)
    qs = [np.ascontiguousarray(q, dtype=np.uint8) for q in queries]
    qptr = (C.c_void_p * max(1, len(qs)))(*[q.ctypes.data for q in qs])
    qlen = np.array([len(q) for q in qs] or [0], dtype=np.int32)
    st = lib.blz_synth_db_fill(C.byref(spec), qptr, qlen.ctypes.data, len(qs), residues.ctypes.data,
                               offsets.ctypes.data, hom.ctypes.data)
    if st:
        raise BlzError(st, "synth_db_fill")
    return residues[: r.value], offsets, hom


# ---------------------------------------------------------------------------------------
# engine
# ---------------------------------------------------------------------------------------
class Result:
    """Per-subject HSP lists of one query: oids (ascending), counts, hsps (subject-major)."""

    def __init__(self, oids, counts, hsps):
        self.oids = oids
        self.counts = counts
        self.hsps = hsps

    def as_dict(self) -> dict:
        out, o = {}, 0
        for oid, c in zip(self.oids.tolist(), self.counts.tolist()):
            out[oid] = self.hsps[o:o + c]
            o += c
        return out


class Engine:
    """One HIP device, one stream, a resident database and prepared queries."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.blz_ctx_create(device, C.byref(h))
        if st:
            raise BlzError(st, "blz_ctx_create: no usable gfx950 device")
        self.ctx = h
        self.device = device
        self.db = None
        self.db_n = 0

    def _err(self, st: int, what: str):
        msg = self.lib.blz_last_error(self.ctx)
        raise BlzError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if self.db:
            self.lib.blz_db_free(self.db)
            self.db = None
        if self.ctx:
            self.lib.blz_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tuning(self, **kw):
        t = Tuning()
        self.lib.blz_ctx_get_tuning(self.ctx, C.byref(t))
        names = {f for f, _ in Tuning._fields_}
        for k, v in kw.items():
            if k not in names:
                raise TypeError(f"unknown tuning field {k!r}")
            setattr(t, k, v)
        st = self.lib.blz_ctx_set_tuning(self.ctx, C.byref(t))
        if st:
            self._err(st, "blz_ctx_set_tuning")

    def get_tuning(self) -> dict:
        t = Tuning()
        self.lib.blz_ctx_get_tuning(self.ctx, C.byref(t))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    @property
    def stream(self) -> int:
        return self.lib.blz_ctx_stream(self.ctx)

    def upload_db(self, residues: np.ndarray, offsets: np.ndarray):
        residues = np.ascontiguousarray(residues, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        if self.db:
            self.lib.blz_db_free(self.db)
            self.db = None
        h = C.c_void_p()
        n = len(offsets) - 1
        st = self.lib.blz_db_upload(self.ctx, residues.ctypes.data, offsets.ctypes.data, n, C.byref(h))
        if st:
            self._err(st, "blz_db_upload")
        self.db = h
        self.db_n = n

    def upload_blastdb(self, bdb: "BlastDB", sort_by_length: bool = True) -> np.ndarray:
        """Makes a BLAST database resident; returns the OID of each uploaded position."""
        if self.db:
            self.lib.blz_db_free(self.db)
            self.db = None
        h = C.c_void_p()
        perm = np.empty(max(1, bdb.num_seqs), np.int64)
        st = self.lib.blz_db_upload_blastdb(self.ctx, bdb.h, int(sort_by_length), C.byref(h), perm.ctypes.data)
        if st:
            self._err(st, "blz_db_upload_blastdb")
        self.db = h
        self.db_n = bdb.num_seqs
        return perm[: bdb.num_seqs]

    def prepare(self, query: np.ndarray, params: Params | None = None):
        query = np.ascontiguousarray(query, dtype=np.uint8)
        h = C.c_void_p()
        st = self.lib.blz_query_prepare(self.ctx, query.ctypes.data, len(query),
                                        C.byref(params) if params is not None else None, C.byref(h))
        if st:
            self._err(st, "blz_query_prepare")
        return h

    def free_query(self, qh):
        self.lib.blz_query_free(qh)

    @staticmethod
    def _sink(n_subj_cap: int, hsp_cap: int):
        oids = np.empty(max(1, n_subj_cap), dtype=np.int64)
        counts = np.empty(max(1, n_subj_cap), dtype=np.int32)
        hsps = np.empty(max(1, hsp_cap), dtype=HSP_DTYPE)
        s = Sink(oids.ctypes.data_as(C.POINTER(C.c_int64)), counts.ctypes.data_as(C.POINTER(C.c_int32)),
                 n_subj_cap, hsps.ctypes.data, hsp_cap, 0, 0)
        return s, (oids, counts, hsps)

    def _fetch(self, qi: int) -> "Result":
        ns, nh = C.c_int64(), C.c_int64()
        st = self.lib.blz_result_sizes(self.ctx, qi, C.byref(ns), C.byref(nh))
        if st:
            self._err(st, "blz_result_sizes")
        sink, (oids, counts, hsps) = self._sink(ns.value, nh.value)
        st = self.lib.blz_result_copy(self.ctx, qi, C.byref(sink))
        if st:
            self._err(st, "blz_result_copy")
        return Result(oids[: ns.value], counts[: ns.value], hsps[: nh.value])

    def hitlist(self, qi: int, stats: SearchStats):
        """Query qi's hit list of the last run (best subjects by e-value, in that order) and the
        best e-value of each listed subject."""
        ns, nh = C.c_int64(), C.c_int64()
        st = self.lib.blz_result_sizes(self.ctx, qi, C.byref(ns), C.byref(nh))
        if st:
            self._err(st, "blz_result_sizes")
        sink, (oids, counts, hsps) = self._sink(ns.value, nh.value)
        ev = np.empty(max(1, ns.value), dtype=np.float64)
        st = self.lib.blz_result_hitlist(self.ctx, qi, C.byref(stats), C.byref(sink), ev.ctypes.data)
        if st:
            self._err(st, "blz_result_hitlist")
        k, h = sink.n_subjects, sink.n_hsps
        return Result(oids[:k], counts[:k], hsps[:h]), ev[:k]

    def overflow_pairs(self) -> list[tuple[int, int]]:
        """(query index, subject) pairs the last run handed back (BLZ_EOVERFLOW)."""
        n = C.c_int64()
        self.lib.blz_overflow_drain(self.ctx, None, 0, C.byref(n))
        buf = np.empty(max(1, n.value), dtype=np.int64)
        st = self.lib.blz_overflow_drain(self.ctx, buf.ctypes.data, n.value, C.byref(n))
        if st:
            self._err(st, "blz_overflow_drain")
        return [(int(v >> 40), int(v & ((1 << 40) - 1))) for v in buf[: n.value]]

    def run(self, qh, allow_overflow: bool = False):
        """Prelim stage of one prepared query over the resident database -> (Result, stats)."""
        stats = RunStats()
        st = self.lib.blz_prelim_run(self.ctx, qh, self.db, None, C.byref(stats))
        if st and not (st == 5 and allow_overflow):
            self._err(st, "blz_prelim_run")
        return self._fetch(0), stats.as_dict()

    def run_batch(self, qhs, allow_overflow: bool = False, fetch: bool = True):
        """Prelim stage of several prepared queries in one batched run -> ([Result], stats)."""
        n = len(qhs)
        stats = RunStats()
        arr = (C.c_void_p * n)(*[q.value if isinstance(q, C.c_void_p) else q for q in qhs])
        st = self.lib.blz_prelim_run_batch(self.ctx, arr, n, self.db, None, C.byref(stats))
        if st and not (st == 5 and allow_overflow):
            self._err(st, "blz_prelim_run_batch")
        return ([self._fetch(i) for i in range(n)] if fetch else None), stats.as_dict()

    def traceback(self, qhs, lists, params: TbParams | None = None, cutoffs=None):
        """Traceback of the listed pairs: lists[i] is a Result (e.g. a hit list) of prepared query
        qhs[i]; cutoffs[i] (optional) its score cutoff in place of params.cutoff.  Returns a
        TbResult (per query, {oid: (traced HSPs as TB_DTYPE in score order, their runs)}, op_first
        indexing that pair's runs) and the call's stats."""
        nq = len(qhs)
        per = [len(r.oids) for r in lists]
        n = int(sum(per))
        pa = np.empty(n, TB_PAIR_DTYPE)
        if n:
            pa["query"] = np.repeat(np.arange(len(lists), dtype=np.int32), per)
            pa["n_hsps"] = np.concatenate([np.asarray(r.counts, np.int32) for r in lists])
            pa["oid"] = np.concatenate([np.asarray(r.oids, np.int64) for r in lists])
        hs = np.ascontiguousarray(np.concatenate([np.asarray(r.hsps, dtype=HSP_DTYPE) for r in lists])
                                  if lists else np.empty(0, HSP_DTYPE))
        arr = (C.c_void_p * max(1, nq))(*[q.value if isinstance(q, C.c_void_p) else q for q in qhs])
        p = params if params is not None else tb_params()
        stats = TbStats()
        cut = None if cutoffs is None else np.ascontiguousarray(cutoffs, dtype=np.int32)
        st = self.lib.blz_traceback(self.ctx, self.db, arr, nq, None if cut is None else cut.ctypes.data,
                                    pa.ctypes.data, n, hs.ctypes.data, C.byref(p), C.byref(stats))
        if st:
            self._err(st, "blz_traceback")
        n_p, n_h, n_r = C.c_int64(), C.c_int64(), C.c_int64()
        self.lib.blz_traceback_sizes(self.ctx, C.byref(n_p), C.byref(n_h), C.byref(n_r))
        counts = np.empty(max(1, n_p.value), np.int32)
        tb = np.empty(max(1, n_h.value), TB_DTYPE)
        runs = np.empty(max(1, n_r.value), np.uint32)
        st = self.lib.blz_traceback_copy(self.ctx, counts.ctypes.data, n_p.value, tb.ctypes.data, n_h.value,
                                         runs.ctypes.data, n_r.value)
        if st:
            self._err(st, "blz_traceback_copy")
        return TbResult(nq, pa, counts[: n_p.value], tb[: n_h.value], runs[: n_r.value]), stats.as_dict()

class TbResult:
    """Result of one traceback call: pair i (in call order) has counts[i] traced HSPs; result[qi]
    maps a subject oid of query qi to (its traced HSPs, their runs) with op_first rebased."""

    def __init__(self, nq, pairs, counts, tb, runs):
        self.pairs, self.counts, self.tb, self.runs = pairs, counts, tb, runs
        self.first = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])
        self._views = [None] * nq

    def pair(self, i: int):
        t = self.tb[self.first[i]:self.first[i + 1]].copy()
        if len(t) == 0:
            return t, np.empty(0, np.uint32)
        r0 = int(t["op_first"][0])
        r1 = int(t["op_first"][-1] + t["op_count"][-1])
        t["op_first"] -= r0
        return t, self.runs[r0:r1].copy()

    def __len__(self):
        return len(self._views)

    def __getitem__(self, qi: int) -> "TbQueryView":
        if self._views[qi] is None:
            idx = np.nonzero(self.pairs["query"] == qi)[0]
            self._views[qi] = TbQueryView(self, dict(zip(self.pairs["oid"][idx].tolist(), idx.tolist())))
        return self._views[qi]


class TbQueryView:
    def __init__(self, res: TbResult, index: dict):
        self.res, self.index = res, index

    def __getitem__(self, oid):
        return self.res.pair(self.index[oid])

    def __contains__(self, oid):
        return oid in self.index

    def __len__(self):
        return len(self.index)

    def __iter__(self):
        return iter(self.index)

    def keys(self):
        return self.index.keys()

    def items(self):
        return ((oid, self.res.pair(i)) for oid, i in self.index.items())


def tb_params(**kw) -> TbParams:
    p = TbParams()
    load_library().blz_tb_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Server:
    """Request loop over an engine's resident database (blz_server_*): submit queries from any
    thread, wait for each, fetch its lists."""

    def __init__(self, engine: "Engine", slots: int = 0, max_batch: int = 0, max_batch_letters: int = 0,
                 tuning: dict | None = None, max_batch_work: int = 0, overlap_ms: float = 0.0, linger_us: int = 0):
        self.lib = engine.lib
        cfg = ServerConfig()
        self.lib.blz_server_default_config(C.byref(cfg))
        if slots:
            cfg.slots = slots
        if max_batch:
            cfg.max_batch = max_batch
        if max_batch_letters:
            cfg.max_batch_letters = max_batch_letters
        if max_batch_work:
            cfg.max_batch_work = max_batch_work
        cfg.overlap_ms = overlap_ms
        cfg.linger_us = linger_us
        t = None
        if tuning:
            t = Tuning()
            engine.lib.blz_ctx_get_tuning(engine.ctx, C.byref(t))
            for k, v in tuning.items():
                setattr(t, k, v)
        h = C.c_void_p()
        st = self.lib.blz_server_create(engine.device, engine.db, C.byref(cfg), C.byref(t) if t else None,
                                        C.byref(h))
        if st:
            raise BlzError(st, "blz_server_create")
        self.s = h

    def submit(self, query: np.ndarray, params: Params | None = None) -> int:
        query = np.ascontiguousarray(query, dtype=np.uint8)
        tk = C.c_uint64()
        st = self.lib.blz_server_submit(self.s, query.ctypes.data, len(query),
                                        C.byref(params) if params is not None else None, C.byref(tk))
        if st:
            raise BlzError(st, "blz_server_submit")
        return tk.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> dict:
        info = RequestInfo()
        st = self.lib.blz_server_wait(self.s, ticket, timeout_ms, C.byref(info))
        if st:
            raise BlzError(st, "blz_server_wait")
        return {k: getattr(info, k) for k, _ in RequestInfo._fields_}

    def result(self, ticket: int) -> "Result":
        info = self.wait(ticket)
        sink, (oids, counts, hsps) = Engine._sink(info["n_subjects"], info["n_hsps"])
        st = self.lib.blz_server_copy(self.s, ticket, C.byref(sink))
        if st:
            raise BlzError(st, "blz_server_copy")
        return Result(oids[: info["n_subjects"]], counts[: info["n_subjects"]], hsps[: info["n_hsps"]])

    def overflow(self, ticket: int) -> list[int]:
        """Subjects the request's batch handed back."""
        n = C.c_int64()
        self.lib.blz_server_overflow(self.s, ticket, None, 0, C.byref(n))
        buf = np.empty(max(1, n.value), dtype=np.int64)
        st = self.lib.blz_server_overflow(self.s, ticket, buf.ctypes.data, n.value, C.byref(n))
        if st:
            raise BlzError(st, "blz_server_overflow")
        return [int(v) for v in buf[: n.value]]

    def release(self, ticket: int) -> None:
        self.lib.blz_server_release(self.s, ticket)

    def stats(self) -> dict:
        s = ServerStats()
        self.lib.blz_server_get_stats(self.s, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self) -> None:
        if self.s:
            self.lib.blz_server_destroy(self.s)
            self.s = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
