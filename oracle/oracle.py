"""ctypes access to the CPU oracle (oracle/liboracle.so) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only as the checker or the timed CPU restatement; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB = _HERE / "liboracle.so"

HSP_DTYPE = np.dtype(
    [("score", "<i4"), ("q_off", "<i4"), ("q_end", "<i4"), ("s_off", "<i4"), ("s_end", "<i4"),
     ("q_gs", "<i4"), ("s_gs", "<i4")]
)
INIT_DTYPE = np.dtype(
    [("q_hit", "<i4"), ("s_hit", "<i4"), ("q_start", "<i4"), ("s_start", "<i4"), ("len", "<i4"),
     ("score", "<i4")]
)


class Params(C.Structure):
    _fields_ = [("word_threshold", C.c_int32), ("window", C.c_int32), ("x_ungapped", C.c_int32),
                ("x_gapped", C.c_int32), ("gap_open", C.c_int32), ("gap_extend", C.c_int32),
                ("ungapped_cutoff", C.c_int32), ("gapped_cutoff", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("word_hits", C.c_uint64), ("ungapped_ext", C.c_uint64), ("init_hits", C.c_uint64),
                ("gapped_ext", C.c_uint64), ("dp_cells", C.c_uint64), ("hsps", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Lookup(C.Structure):
    _fields_ = [("prefix", C.POINTER(C.c_uint32)), ("entries", C.POINTER(C.c_int32)),
                ("qlen", C.c_int32)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    L = C.CDLL(str(LIB))
    vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
    L.ora_default_params.argtypes = [C.POINTER(Params)]
    L.ora_matrix.restype = C.POINTER(C.c_int32)
    L.ora_neighbors.argtypes = [i32, i32, vp]
    L.ora_neighbors.restype = i32
    L.ora_lookup_build.argtypes = [vp, i32, i32]
    L.ora_lookup_build.restype = C.POINTER(Lookup)
    L.ora_lookup_free.argtypes = [C.POINTER(Lookup)]
    L.ora_extend_left.argtypes = [vp, vp, i32, i32, i32, i32, C.POINTER(i32)]
    L.ora_extend_left.restype = i32
    L.ora_extend_right.argtypes = [vp, i32, vp, i32, i32, i32, i32, i32, C.POINTER(i32), C.POINTER(i32)]
    L.ora_extend_right.restype = i32
    L.ora_word_finder.argtypes = [vp, i32, C.POINTER(Lookup), vp, i32, C.POINTER(Params), vp, i32,
                                  C.POINTER(Stats)]
    L.ora_word_finder.restype = i32
    L.ora_gapped_start.argtypes = [vp, vp, i32, i32, i32]
    L.ora_gapped_start.restype = i32
    L.ora_semi_gapped.argtypes = [vp, vp, i32, i32, i32, i32, i32, i32, C.POINTER(i32), C.POINTER(i32),
                                  C.POINTER(C.c_uint64)]
    L.ora_semi_gapped.restype = i32
    L.ora_prelim_subject.argtypes = [vp, i32, C.POINTER(Lookup), vp, i32, C.POINTER(Params), vp, i32,
                                     C.POINTER(Stats)]
    L.ora_prelim_subject.restype = i32
    L.ora_prelim_db.argtypes = [vp, i32, vp, vp, i64, C.POINTER(Params), i32, vp, vp, i64,
                                C.POINTER(Stats)]
    L.ora_prelim_db.restype = i64
    L.ora_gapped_stage.argtypes = [vp, i32, vp, i32, vp, i32, C.POINTER(Params), vp, i32, C.POINTER(Stats)]
    L.ora_gapped_stage.restype = i32
    L.ora_purge.argtypes = [vp, i32]
    L.ora_purge.restype = i32
    L.ora_sort_hsps.argtypes = [vp, i32]
    L.ora_align_ex.argtypes = [vp, vp, i32, i32, i32, i32, i32, i32, C.POINTER(i32), C.POINTER(i32), vp,
                               C.POINTER(i32), C.POINTER(C.c_uint64)]
    L.ora_align_ex.restype = i32
    L.ora_traceback_subject.argtypes = [vp, i32, vp, i32, vp, i32, i32, i32, i32, i32, vp, i32, vp, i32,
                                        C.POINTER(i32), C.POINTER(C.c_uint64)]
    L.ora_traceback_subject.restype = i32
    _lib = L
    return L


def params(**kw) -> Params:
    p = Params()
    lib().ora_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def matrix() -> np.ndarray:
    ptr = lib().ora_matrix()
    return np.ctypeslib.as_array(ptr, shape=(28 * 28,)).reshape(28, 28).copy()


def neighbors(word: int, threshold: int = 11) -> np.ndarray:
    out = np.empty(28 ** 3 + 1, dtype=np.int32)
    n = lib().ora_neighbors(word, threshold, out.ctypes.data)
    return np.sort(out[:n])


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint8)


class LookupTable:
    def __init__(self, query, threshold: int = 11):
        self.query = _u8(query)
        self.ptr = lib().ora_lookup_build(self.query.ctypes.data, len(self.query), threshold)
        total = self.ptr.contents.prefix[28 ** 3]
        self.prefix = np.ctypeslib.as_array(self.ptr.contents.prefix, shape=(28 ** 3 + 1,)).copy()
        self.entries = (np.ctypeslib.as_array(self.ptr.contents.entries, shape=(max(total, 1),))[:total].copy())

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().ora_lookup_free(self.ptr)
            self.ptr = None


def extend_left(q, s, q_off, s_off, dropoff=16, maxscore=0):
    q, s = _u8(q), _u8(s)
    n = C.c_int()
    sc = lib().ora_extend_left(q.ctypes.data, s.ctypes.data, q_off, s_off, dropoff, maxscore, C.byref(n))
    return sc, n.value


def extend_right(q, s, q_off, s_off, dropoff=16, maxscore=0):
    q, s = _u8(q), _u8(s)
    n, last = C.c_int(), C.c_int()
    sc = lib().ora_extend_right(q.ctypes.data, len(q), s.ctypes.data, len(s), q_off, s_off, dropoff, maxscore,
                                C.byref(n), C.byref(last))
    return sc, n.value, last.value


def semi_gapped(A, B, x_dropoff=38, gap_open=11, gap_extend=1, reverse=False, M=None, N=None):
    A, B = _u8(A), _u8(B)
    if M is None:
        M = len(A) if reverse else len(A) - 1
    if N is None:
        N = len(B) if reverse else len(B) - 1
    a, b, cells = C.c_int(), C.c_int(), C.c_uint64(0)
    sc = lib().ora_semi_gapped(A.ctypes.data, B.ctypes.data, M, N, x_dropoff, gap_open, gap_extend,
                               1 if reverse else 0, C.byref(a), C.byref(b), C.byref(cells))
    return sc, a.value, b.value, cells.value


def gapped_start(q, s, q_start, s_start, length):
    q, s = _u8(q), _u8(s)
    return lib().ora_gapped_start(q.ctypes.data, s.ctypes.data, q_start, s_start, length)


def word_finder(query, subject, p: Params | None = None, lookup: LookupTable | None = None):
    p = p or params()
    query, subject = _u8(query), _u8(subject)
    lk = lookup or LookupTable(query, p.word_threshold)
    st = Stats()
    cap = 4096
    while True:
        out = np.empty(cap, dtype=INIT_DTYPE)
        n = lib().ora_word_finder(query.ctypes.data, len(query), lk.ptr, subject.ctypes.data, len(subject),
                                  C.byref(p), out.ctypes.data, cap, C.byref(st))
        if n <= cap:
            return out[:n], st.as_dict()
        cap = n
        st = Stats()


def prelim_subject(query, subject, p: Params | None = None, lookup: LookupTable | None = None):
    p = p or params()
    query, subject = _u8(query), _u8(subject)
    lk = lookup or LookupTable(query, p.word_threshold)
    cap = 1024
    while True:
        st = Stats()
        out = np.empty(cap, dtype=HSP_DTYPE)
        n = lib().ora_prelim_subject(query.ctypes.data, len(query), lk.ptr, subject.ctypes.data, len(subject),
                                     C.byref(p), out.ctypes.data, cap, C.byref(st))
        if n <= cap:
            return out[:n], st.as_dict()
        cap = n


def gapped_stage(query, subject, init_hits, p: Params | None = None):
    """Gapped stage over given init hits (records or (q_start, s_start, len, score) tuples)."""
    p = p or params()
    query, subject = _u8(query), _u8(subject)
    ih = np.zeros(len(init_hits), dtype=INIT_DTYPE)
    for k, h in enumerate(init_hits):
        if isinstance(h, tuple):
            ih[k]["q_start"], ih[k]["s_start"], ih[k]["len"], ih[k]["score"] = h
            ih[k]["q_hit"], ih[k]["s_hit"] = h[0], h[1]
        else:
            ih[k] = h
    cap = max(1, len(init_hits))
    st = Stats()
    out = np.empty(cap, dtype=HSP_DTYPE)
    n = lib().ora_gapped_stage(query.ctypes.data, len(query), subject.ctypes.data, len(subject), ih.ctypes.data,
                               len(ih), C.byref(p), out.ctypes.data, cap, C.byref(st))
    return out[:n], st.as_dict()


def hsp_records(rows) -> np.ndarray:
    """(score, q_off, q_end, s_off, s_end, q_gs, s_gs) tuples -> HSP records."""
    return np.array([tuple(r) for r in rows], dtype=HSP_DTYPE)


def purge(hsps) -> np.ndarray:
    h = np.ascontiguousarray(hsps, dtype=HSP_DTYPE).copy()
    n = lib().ora_purge(h.ctypes.data, len(h))
    return h[:n]


def sort_hsps(hsps) -> np.ndarray:
    h = np.ascontiguousarray(hsps, dtype=HSP_DTYPE).copy()
    lib().ora_sort_hsps(h.ctypes.data, len(h))
    return h


TB_DTYPE = np.dtype(
    [("score", "<i4"), ("q_off", "<i4"), ("q_end", "<i4"), ("s_off", "<i4"), ("s_end", "<i4"), ("q_gs", "<i4"),
     ("s_gs", "<i4"), ("num_ident", "<i4"), ("align_len", "<i4"), ("gaps", "<i4"), ("gap_opens", "<i4"),
     ("op_first", "<i4"), ("op_count", "<i4")]
)
OP_SUB, OP_GAP_Q, OP_GAP_S = 0, 1, 2


def align_ex(A, B, x_dropoff, gap_open=11, gap_extend=1, reverse=False, M=None, N=None):
    """(score, a_off, b_off, ops from the best cell back to the origin, cells) of one direction."""
    A, B = _u8(A), _u8(B)
    if M is None:
        M = len(A) if reverse else len(A) - 1
    if N is None:
        N = len(B) if reverse else len(B) - 1
    a, b, n, cells = C.c_int(), C.c_int(), C.c_int(), C.c_uint64(0)
    ops = np.empty(M + N + 8, np.uint8)
    sc = lib().ora_align_ex(A.ctypes.data, B.ctypes.data, M, N, x_dropoff, gap_open, gap_extend, 1 if reverse else 0,
                            C.byref(a), C.byref(b), ops.ctypes.data, C.byref(n), C.byref(cells))
    return sc, a.value, b.value, ops[: n.value].copy(), cells.value


def traceback_subject(query, subject, prelim_hsps, x_final=64, gap_open=11, gap_extend=1, cutoff=0,
                      return_cells=False):
    """Traced HSPs (TB_DTYPE, final order) of one subject, and their runs (op << 28 | length);
    with return_cells also the DP cells the subject's alignments evaluated."""
    query, subject = _u8(query), _u8(subject)
    pre = np.ascontiguousarray(prelim_hsps, dtype=HSP_DTYPE)
    cap = max(1, len(pre))
    out = np.empty(cap, TB_DTYPE)
    runs_cap = 4 * (len(query) + len(subject)) * cap + 16
    runs = np.empty(runs_cap, np.uint32)
    nr, cells = C.c_int(), C.c_uint64(0)
    n = lib().ora_traceback_subject(query.ctypes.data, len(query), subject.ctypes.data, len(subject), pre.ctypes.data,
                                    len(pre), x_final, gap_open, gap_extend, cutoff, out.ctypes.data, cap,
                                    runs.ctypes.data, runs_cap, C.byref(nr), C.byref(cells))
    assert nr.value >= 0
    if return_cells:
        return out[:n], runs[: nr.value].copy(), cells.value
    return out[:n], runs[: nr.value].copy()


def prelim_db(query, residues, offsets, p: Params | None = None, nthreads: int = 1, hsp_cap: int = 1 << 20):
    """Returns (counts int32[n], hsps subject-major, stats dict)."""
    p = p or params()
    query, residues = _u8(query), _u8(residues)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = len(offsets) - 1
    counts = np.zeros(max(n, 1), dtype=np.int32)
    while True:
        st = Stats()
        out = np.empty(max(hsp_cap, 1), dtype=HSP_DTYPE)
        tot = lib().ora_prelim_db(query.ctypes.data, len(query), residues.ctypes.data, offsets.ctypes.data, n,
                                  C.byref(p), nthreads, counts.ctypes.data, out.ctypes.data, hsp_cap, C.byref(st))
        if tot >= 0:
            return counts[:n], out[:tot], st.as_dict()
        hsp_cap = int(counts[:n].sum())


def as_lists(counts, hsps) -> dict:
    """{oid: hsp array} for subjects with at least one HSP."""
    out, o = {}, 0
    for oid, c in enumerate(counts.tolist()):
        if c:
            out[oid] = hsps[o:o + c]
            o += c
    return out
