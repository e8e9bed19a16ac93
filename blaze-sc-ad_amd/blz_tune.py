"""Autotuner for the batched prelim stage.

Picks the launch shape (waves per seed workgroup), the query-size bin edge (diagonal-ring
length of the first bin; bins double from there), the pack length for short queries, where the
lookup tables live (global memory, entries in LDS, or prefix and entries in LDS: the LDS tile of
the seed kernel), whether short subjects share a wave's scan (subject groups: the engine's
subject-length rule, in place of fixed length bins), the gapped-stage routing and its lanes per
DP row, by timing batched runs of a sample batch over the resident database, one field at a
time (coordinate descent from the built-in choices).  Times are device times of the whole run
(blz_run_stats.ms_total), so host result handling does not enter the choice.

No reference counterpart: the reference hard-codes 64 threads per subject and its bins
(BLAZE_gpu/gpuBLAZE.cu:1639; BLAZE_gpu/util.hpp:124-131).
"""
from __future__ import annotations

from typing import Callable

CANDIDATES: dict[str, list[int]] = {
    "seed_waves": [0, 6, 8, 12, 16],
    "bin_ring": [0, 800, 1600, 2400],
    "pack_len": [0, 600, 1400],
    "seed_table": [0, 1, 2, 3],
    "subject_groups": [0, 1],
    "gapped_wave_classes": [0, 8],
    "gapped_lanes": [0, 64],
    "gapped_width": [0, 16],
}


def time_batch(engine, qhs, tuning: dict, repeats: int = 1) -> float:
    """Device milliseconds of one batched run under `tuning` (best of `repeats`)."""
    engine.set_tuning(**tuning)
    best = float("inf")
    for _ in range(repeats):
        _, st = engine.run_batch(qhs, allow_overflow=True, fetch=False)
        best = min(best, st["ms_total"])
    return best


def tune(engine, qhs, candidates: dict[str, list[int]] | None = None, repeats: int = 2,
         margin: float = 0.05, log: Callable[[str], None] | None = None) -> tuple[dict, float]:
    """Returns (best tuning, its device ms) and leaves the engine set to it.  Each candidate is
    timed `repeats` times (best kept) and replaces the current choice only when it is faster by
    more than `margin`: run-to-run noise on a sample batch is a few percent, and a choice that
    wins on the sample by less than that has lost on the full batch.  Graph capture is switched
    off while tuning (every candidate is a new launch sequence)."""
    candidates = candidates or CANDIDATES
    start = {**engine.get_tuning(), "graph": 1}
    best = dict(start)
    best_ms = time_batch(engine, qhs, best, repeats)
    if log:
        log(f"tune: built-in {best_ms:.2f} ms")
    for key, values in candidates.items():
        for v in values:
            if v == best[key]:
                continue
            trial = {**best, key: v}
            try:
                ms = time_batch(engine, qhs, trial, repeats)
            except RuntimeError as e:  # a choice this batch cannot run (e.g. tables past the LDS)
                if log:
                    log(f"tune: {key}={v} refused ({e})")
                continue
            if log:
                log(f"tune: {key}={v} {ms:.2f} ms")
            if ms < best_ms * (1.0 - margin):  # keep a change only when it is clearly faster
                best, best_ms = trial, ms
    best["graph"] = 0  # capture again once the choice is made
    engine.set_tuning(**best)
    return best, best_ms
