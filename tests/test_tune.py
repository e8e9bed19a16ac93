"""The autotuner's search (blaze-sc-ad_amd/blz_tune.py) on a stand-in engine, on the CPU: a
candidate replaces the current choice only when it is faster by more than the margin, the search
continues from the kept choice, graph capture is off while timing and on again at the end, and the
engine is left set to the result."""
import blz_tune


class FakeEngine:
    """Device time as a function of the tuning: a table of per-field costs."""

    def __init__(self, cost):
        self.cost = cost
        self.tuning = {k: 0 for k in blz_tune.CANDIDATES}
        self.tuning["graph"] = 0
        self.graph_while_timing = []

    def get_tuning(self):
        return dict(self.tuning)

    def set_tuning(self, **kw):
        self.tuning.update(kw)

    def run_batch(self, qhs, allow_overflow=False, fetch=True):
        self.graph_while_timing.append(self.tuning["graph"])
        ms = 100.0 + sum(self.cost.get((k, v), 0.0) for k, v in self.tuning.items())
        return None, {"ms_total": ms}


def test_keeps_only_clear_wins_and_leaves_engine_set():
    cost = {("seed_waves", 8): -10.0,   # 10% faster: kept
            ("bin_ring", 800): -3.0,    # 3%: within the margin, not kept
            ("seed_table", 2): -20.0,   # kept, from the already improved choice
            ("gapped_lanes", 64): +5.0}
    eng = FakeEngine(cost)
    best, ms = blz_tune.tune(eng, qhs=[None], repeats=1, margin=0.05)
    assert best["seed_waves"] == 8 and best["bin_ring"] == 0 and best["seed_table"] == 2
    assert best["gapped_lanes"] == 0
    assert ms == 100.0 - 10.0 - 20.0
    assert set(eng.graph_while_timing) == {1}          # captured graphs off while timing
    assert eng.tuning["graph"] == 0 and eng.tuning["seed_table"] == 2  # set to the result


def test_candidates_cover_launch_shape_bins_tiles_and_routing():
    for k in ("seed_waves", "bin_ring", "pack_len", "seed_table", "subject_groups", "gapped_wave_classes",
              "gapped_lanes", "gapped_width"):
        assert k in blz_tune.CANDIDATES and blz_tune.CANDIDATES[k][0] == 0  # built-in first


def test_a_refused_candidate_is_skipped():
    class Refusing(FakeEngine):
        def run_batch(self, qhs, allow_overflow=False, fetch=True):
            if self.tuning["seed_table"] == 3:
                raise RuntimeError("tables do not fit the LDS")
            return super().run_batch(qhs, allow_overflow, fetch)

    eng = Refusing({("seed_table", 3): -50.0, ("pack_len", 600): -30.0})
    best, ms = blz_tune.tune(eng, qhs=[None], repeats=1)
    assert best["seed_table"] == 0 and best["pack_len"] == 600 and ms == 70.0
