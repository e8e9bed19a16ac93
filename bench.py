"""Benchmark: BLASTP prelim stage over a resident synthetic database.  Prints one JSON line.

--workload scan (default): one 500-aa query against a 2,000,000-subject DB.  A step is one full
  prelim search of the query (seed scan, two-hit, ungapped X-drop, gapped stage, HSP lists back
  on the host).
--workload batch: 256 queries (log-uniform 50-3,000 aa) against the same DB in one batched run
  per step (packs of short queries, one persistent launch per query-size bin, replayed as a
  hipGraph), against the same 256 queries run one query per launch.
--workload stress: 64 queries of 1,000-2,000 aa against 500,000 long subjects (median 800, up
  to 35,000) with 20% homologs at 30-45% identity and low-complexity repeat blocks: GCUPS of the
  semi-gapped stage, and bit-exact lists on a sample against the CPU restatement.
--workload traceback: 256 queries (log-uniform 50-3,000 aa) with ~500 homologs each among 2,000,000
  subjects; their hit lists (expect 10, 500 subjects each) traced (gapped alignment with
  traceback, edit scripts back on the host) in one call per step: pairs/s, traceback GCUPS, a
  sample bit-exact against the CPU restatement.
--workload serve: Poisson query streams (lengths log-normal, median 350, sigma 0.8, clipped
  30-4,000) at several rates against a resident DB (--subjects, 1,000,000 by default; 40,000,000
  is the full serving workload), through the serving loop with continuous batching (--slots
  contexts) and through the same loop with one query per launch: p50/p99 latency, residues/s.

    python bench.py [--workload scan|batch|stress|serve] [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "blaze-sc-ad_amd"))

import blz  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "DB residues/s + semi-gapped GCUPS, bit-exact HSPs; p50/p99 query latency"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(n_subjects: int, qlen: int, seed: int):
    q = blz.synth_query(seed, qlen)
    spec = blz.synth_spec(seed=seed, n_subjects=n_subjects, len_median=300, len_sigma=0.8, len_min=30,
                          len_max=5000, homolog_frac=0.01, ident_min=0.35, ident_max=0.70)
    res, offs, hom = blz.synth_db(spec, [q])
    return q, res, offs, hom


def host_cpu() -> dict:
    """nproc (the cores this process may use) and the CPU model name."""
    model = "unknown"
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": len(os.sched_getaffinity(0)), "model": model}


def _timed_runs(fn, runs: int):
    out = []
    for _ in range(runs):
        t = time.perf_counter()
        r = fn()
        out.append(time.perf_counter() - t)
    return float(np.median(out)), r


def cpu_baseline(q, res, offs, got: blz.Result, budget_s: float, runs: int = 5):
    """Times the CPU restatement (oracle) on the host's cores: all of them (nproc threads) and one
    thread, each the median of `runs` runs after one warm-up, on a stride sample of the same
    database sized so that the whole measurement takes about budget_s; the query's lookup-table
    build is inside every run.  The all-core sample's lists are checked against the GPU's."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    host = host_cpu()
    threads = host["nproc"]
    n = len(offs) - 1

    def sample(stride):
        idx = np.arange(0, n, max(1, stride))
        return (idx, *_gather(res, offs, idx))

    # probe the single-thread rate, then size both samples: half the budget each
    idx, sub_res, sub_offs = sample(max(1, n // 2000))
    t = time.perf_counter()
    oracle.prelim_db(q, sub_res, sub_offs, nthreads=1)
    rate1 = float(sub_offs[-1]) / max(1e-6, time.perf_counter() - t)
    per_run = budget_s / 2 / (runs + 1)

    def stride_for(rate):
        return max(1, int(np.ceil(res.size / max(1.0, rate * per_run))))

    idx1, r1, o1 = sample(stride_for(rate1))
    oracle.prelim_db(q, r1, o1, nthreads=1)
    t1, _ = _timed_runs(lambda: oracle.prelim_db(q, r1, o1, nthreads=1), runs)
    idxa, ra, oa = sample(stride_for(rate1 * threads * 0.8))
    oracle.prelim_db(q, ra, oa, nthreads=threads)
    ta, (counts, hsps, _) = _timed_runs(lambda: oracle.prelim_db(q, ra, oa, nthreads=threads), runs)
    want = oracle.as_lists(counts, hsps)
    have_all = got.as_dict()
    have = {k: have_all[int(idxa[k])] for k in range(len(idxa)) if int(idxa[k]) in have_all}
    exact = sorted(want) == sorted(have) and all(np.array_equal(want[k], have[k]) for k in want)

    def desc(idx, o, dt, th):
        frac = "the whole DB" if len(idx) == n else f"every {n // max(1, len(idx))}th subject of the same DB"
        return f"{frac}: {len(idx)} subjects, {int(o[-1])} residues, median {dt:.2f} s of {runs} runs on {th} thread(s)"

    return {
        "value": float(oa[-1] / ta), "unit": "residues/s", "cores": threads, "kind": "port",
        "cpu_model": host["model"], "nproc": host["nproc"], "runs": runs,
        "sample": desc(idxa, oa, ta, threads) + "; CPU restatement of the NCBI prelim stage (oracle/), "
                  "query table build included",
        "single_thread": {"value": float(o1[-1] / t1), "unit": "residues/s", "cores": 1,
                          "sample": desc(idx1, o1, t1, 1)},
    }, exact, len(idxa)


def _gather(res, offs, idx):
    lens = (offs[idx + 1] - offs[idx]).astype(np.int64)
    sub_offs = np.zeros(len(idx) + 1, dtype=np.uint64)
    sub_offs[1:] = np.cumsum(lens)
    sub_res = np.concatenate([res[int(offs[i]):int(offs[i + 1])] for i in idx]) if len(idx) else res[:0]
    return sub_res, sub_offs


def copy_calibration(nbytes: int) -> int:
    import torch

    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "launches": 3}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="scan", choices=["scan", "batch", "stress", "serve", "traceback"])
    ap.add_argument("--rates", default="20,100,200", help="serve: arrival rates, queries/s")
    ap.add_argument("--duration", type=float, default=4.0, help="serve: seconds of arrivals per rate")
    ap.add_argument("--slots", type=int, default=2, help="serve: contexts of the batched form")
    ap.add_argument("--autotune", action="store_true", help="batch: tune on a sample batch first")
    ap.add_argument("--queries", type=int, default=0, help="batch/stress: number of queries (0: 256 / 64)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--subjects", type=int, default=2_000_000)
    ap.add_argument("--qlen", type=int, default=500)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", default="", help="JSON of blz_tuning fields, e.g. '{\"seed_table\": 2}'")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (three times): the reference kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (the path shards by database subsets, one per GPU, with no "
              "collective in the hot loop, so one GPU's share is the figure); run with --gpus 1", file=sys.stderr)
        return 2
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "stress":
        return run_stress(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "traceback":
        return run_traceback(args)
    return run_scan(args)


SEED_SOURCES = ["blz_seed.hip", "blz_dev_common.h", "blz_device.h", "blz_common.h"]


def seed_source_hash() -> str:
    import hashlib

    h = hashlib.sha256()
    for name in SEED_SOURCES:
        h.update((ROOT / "blaze-sc-ad_amd" / "csrc" / name).read_bytes())
    return h.hexdigest()[:16]


def valu_peak() -> tuple[float, str]:
    """VALU issue peak in G wave64 instructions/s: each SIMD issues a wave64 VALU instruction per
    4-cycle quad and can dual-issue a second (the SQ_ACTIVE_INST_VALU2 counter counts the quads that
    do), 4 SIMDs per CU, engine clock 2.4 GHz (rocminfo: Max Clock Freq. 2400 MHz on gfx950)."""
    cus = 256
    try:
        import torch

        cus = torch.cuda.get_device_properties(0).multi_processor_count
    except Exception:
        pass
    clk = 2.4
    return cus * 4 * clk / 4 * 2, f"{cus} CUs x 4 SIMDs x {clk} GHz / 4 cycles per wave64 instruction x 2 (dual issue)"


def seed_roofline(seed_bytes: int, seed_ms: float) -> dict:
    """k_seed2's roofline: VALU instruction issue (binding), HBM bytes (secondary).  Instruction
    counts and traffic come from the committed PMC summaries of this workload; they are used only
    when they were measured on the current seed-kernel sources."""
    out = {"kernel": "k_seed2", "bound": "valu", "unit": "G VALU wave-instructions/s", "achieved": None,
           "peak": None, "frac": None, "traffic": None}
    peak, basis = valu_peak()
    out["peak"], out["peak_basis"] = peak, basis
    hbm = seed_bytes / (seed_ms * 1e-3) / 1e9
    out["hbm"] = {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBS,
                  "algorithmic_bytes": seed_bytes}
    src = seed_source_hash()
    issue = ROOT / "profiles" / "k_seed_issue.json"
    if issue.exists():
        prof = json.loads(issue.read_text())
        if prof.get("algorithmic_bytes") == seed_bytes and prof.get("source_sha") == src:
            ach = prof["valu_insts_per_launch"] / (seed_ms * 1e-3) / 1e9
            out.update(achieved=ach, frac=ach / peak, valu_insts_per_launch=prof["valu_insts_per_launch"],
                       valu_busy=prof.get("valu_busy"), valu_issue_slots=prof.get("valu_issue_slots"),
                       salu_insts_per_launch=prof.get("salu_insts_per_launch"),
                       profile=prof.get("source"))
    traffic = ROOT / "profiles" / "k_seed_traffic.json"
    if traffic.exists():
        prof = json.loads(traffic.read_text())
        if prof.get("algorithmic_bytes") == seed_bytes and prof.get("source_sha") == src:
            out["traffic"] = prof.get("bytes_per_launch")
    return out


def run_scan(args) -> int:
    t0 = time.perf_counter()
    q, res, offs, hom = build_workload(args.subjects, args.qlen, args.seed)
    n_sub, n_res = len(offs) - 1, int(offs[-1])
    log(f"workload built: {n_sub} subjects, {n_res} residues, {int((hom >= 0).sum())} homologs "
        f"({time.perf_counter() - t0:.1f} s)")

    with blz.Engine(0) as eng:
        if args.tune:
            eng.set_tuning(**json.loads(args.tune))
        t = time.perf_counter()
        eng.upload_db(res, offs)
        up_s = time.perf_counter() - t
        log(f"DB upload (pinned staging, side stream, checks included): {up_s:.3f} s, {n_res / up_s / 1e9:.2f} GB/s")
        qh = eng.prepare(q)
        for i in range(args.warmup):
            got, st = eng.run(qh)
            log(f"warmup {i}: seed {st['ms_seed']:.2f} ms gapped {st['ms_gapped']:.2f} ms")
        lat, seeds, gapped, cells = [], [], [], []
        t_start = time.perf_counter()
        for i in range(args.steps):
            t = time.perf_counter()
            got, st = eng.run(qh)
            lat.append(time.perf_counter() - t)
            seeds.append(st["ms_seed"])
            gapped.append(st["ms_gapped"])
            cells.append(st["dp_cells"])
        elapsed = time.perf_counter() - t_start
        eng.free_query(qh)
        tuning = eng.get_tuning()
        # the same search as a caller with a new query each time sees it: table build, upload,
        # planning, capture of a new launch sequence (no replay of an earlier one), lists back on
        # the host; three untimed new queries first (the host side of this path settles over the
        # first few: 14.2, 13.0, 12.6, 12.0, then 11.7 ms on one MI355X)
        fresh_lat = []
        for i in range(3 + 5):
            t = time.perf_counter()
            h2 = eng.prepare(q)
            got2, _ = eng.run(h2)
            eng.free_query(h2)
            if i >= 3:
                fresh_lat.append(time.perf_counter() - t)
        fresh = {"ms_per_step": 1000 * float(np.median(fresh_lat)), "steps": 5, "warmup": 3,
                 "value": n_res / float(np.median(fresh_lat)), "lists_equal": _same_lists(got, got2)}
    log(f"timed {args.steps} steps in {elapsed:.3f} s; last stats {st}")
    log(f"fresh query per step: {fresh['ms_per_step']:.2f} ms")

    ms_step = 1000.0 * elapsed / args.steps
    value = n_res * args.steps / elapsed
    seed_ms = float(np.mean(seeds))
    gapped_ms = float(np.mean(gapped))
    # dominant kernel: the seed scan (k_seed2).  Its algorithmic HBM bytes per launch are the
    # residues once plus the u64 offsets; it is bound by VALU instruction issue, priced with the
    # instruction count of its PMC profile on this same workload and build.
    seed_bytes = n_res + 8 * (n_sub + 1)
    roofline = seed_roofline(seed_bytes, seed_ms)

    cpu, exact = None, None
    if not args.no_cpu:
        cpu, exact, nsample = cpu_baseline(q, res, offs, got, args.cpu_budget)
        log(f"cpu baseline {cpu['value']:.3e} residues/s; GPU lists of the {nsample} sampled subjects "
            f"bit-exact: {exact}")

    out = {
        "metric": METRIC, "value": value, "unit": "residues/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "one 500-aa query vs 2,000,000-subject length-sorted DB "
                               "(log-normal median 300, sigma 0.8, clip 30-5000), 1% homologs at 35-70% "
                               "identity, seed 0; one launch per stage",
                   "subjects": n_sub, "residues": n_res, "query_len": args.qlen},
        "gcups": float(np.mean(cells)) / (gapped_ms * 1e-3) / 1e9 if gapped_ms > 0 else None,
        "latency_ms": {"p50": 1000 * float(np.percentile(lat, 50)), "p99": 1000 * float(np.percentile(lat, 99))},
        "kernel_ms": {"k_seed": seed_ms, "k_gapped": gapped_ms},
        "db_upload": {"seconds": up_s, "GB_per_s": n_res / up_s / 1e9},
        "bit_exact_sample": exact,
        "tuning": tuning,
        "replay": "each step runs the same prepared query again: after the first run the launch sequence is a "
                  "captured hipGraph replay; fresh_query times a new query handle and an eager run per step",
        "fresh_query": fresh,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def _same_lists(a: "blz.Result", b: "blz.Result") -> bool:
    return np.array_equal(a.oids, b.oids) and np.array_equal(a.counts, b.counts) and np.array_equal(a.hsps, b.hsps)


def run_batch(args) -> int:
    """256 queries, log-uniform 50-3,000 aa, against the scan workload's DB: batched and replayed
    as a hipGraph versus one query per launch."""
    nq = args.queries or 256
    rng = np.random.default_rng(args.seed)
    lens = np.exp(rng.uniform(np.log(50), np.log(3000), nq)).astype(int)
    queries = [blz.synth_query(1000 + i, int(L)) for i, L in enumerate(lens)]
    t0 = time.perf_counter()
    q0, res, offs, hom = build_workload(args.subjects, args.qlen, args.seed)
    n_sub, n_res = len(offs) - 1, int(offs[-1])
    log(f"workload built: {nq} queries ({int(lens.sum())} letters), {n_sub} subjects, {n_res} residues "
        f"({time.perf_counter() - t0:.1f} s)")
    with blz.Engine(0) as eng:
        if args.tune:
            eng.set_tuning(**json.loads(args.tune))
        eng.upload_db(res, offs)
        t = time.perf_counter()
        qhs = [eng.prepare(q) for q in queries]
        prep_ms = 1000 * (time.perf_counter() - t)
        tuned = None
        if args.autotune:
            import blz_tune
            sample = qhs[:: max(1, nq // 32)]
            t = time.perf_counter()
            tuned, tuned_ms = blz_tune.tune(eng, sample, log=log)
            log(f"autotune on {len(sample)} queries: {tuned} ({tuned_ms:.1f} ms; {time.perf_counter() - t:.1f} s)")
        for i in range(args.warmup):  # the first run captures the graph
            got, st = eng.run_batch(qhs)
            log(f"warmup {i}: {st['ms_total']:.1f} ms device, {st['seed_launches']} seed launches")
        lat = []
        for i in range(args.steps):
            t = time.perf_counter()
            got, st = eng.run_batch(qhs)
            lat.append(time.perf_counter() - t)
            log(f"step {i}: {1000 * lat[-1]:.1f} ms ({st['ms_total']:.1f} ms device: seed {st['ms_seed']:.1f}, "
                f"gapped {st['ms_gapped']:.1f})")
        tuning = eng.get_tuning()
        # the plain form: one query per launch, no packing, no graph
        eng.set_tuning(**{**tuning, "query_packing": 1, "graph": 1})
        t = time.perf_counter()
        plain, plain_dev = [], 0.0
        for qh in qhs:
            r, s1 = eng.run(qh)
            plain.append(r)
            plain_dev += s1["ms_total"]
        plain_s = time.perf_counter() - t
        exact = all(_same_lists(a, b) for a, b in zip(got, plain))
        for qh in qhs:
            eng.free_query(qh)
    step_s = float(np.mean(lat))
    value = n_res * nq / step_s
    plain_value = n_res * nq / plain_s
    log(f"batched {value:.3e} residues/s, plain {plain_value:.3e}; lists identical: {exact}")
    out = {
        "metric": METRIC, "value": value, "unit": "residues/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000 * step_s, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{nq} queries, log-uniform 50-3,000 aa (seed {args.seed}), vs the 2,000,000-subject "
                               "DB of the scan workload; one batched run per step, replayed as a hipGraph",
                   "queries": nq, "query_letters": int(lens.sum()), "subjects": n_sub, "residues": n_res},
        "device_ms_per_step": st["ms_total"], "seed_launches": st["seed_launches"], "prepare_ms": prep_ms,
        "plain": {"value": plain_value, "ms": 1000 * plain_s, "device_ms": plain_dev,
                  "form": "one blz_prelim_run per query, no packing, no graph"},
        "speedup_vs_plain": value / plain_value, "bit_exact_vs_plain": exact,
        "latency_ms": {"p50": 1000 * float(np.percentile(lat, 50)), "p99": 1000 * float(np.percentile(lat, 99))},
        "tuning": tuning, "autotuned": tuned is not None, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_stress(args) -> int:
    """64 queries of 1,000-2,000 aa against 500,000 long subjects with repeats: GCUPS of the
    semi-gapped stage and bit-exact lists on sampled queries against the CPU restatement."""
    nq = args.queries or 64
    rng = np.random.default_rng(args.seed)
    lens = rng.integers(1000, 2001, nq)
    queries = [blz.synth_query(2000 + i, int(L)) for i, L in enumerate(lens)]
    spec = blz.synth_spec(seed=args.seed, n_subjects=args.subjects if args.subjects != 2_000_000 else 500_000,
                          len_median=800, len_sigma=0.8, len_min=30, len_max=35000, homolog_frac=0.2,
                          ident_min=0.30, ident_max=0.45, repeat_frac=0.05)
    t0 = time.perf_counter()
    res, offs, hom = blz.synth_db(spec, queries)
    n_sub, n_res = len(offs) - 1, int(offs[-1])
    log(f"workload built: {nq} queries, {n_sub} subjects, {n_res} residues, {int((hom >= 0).sum())} homologs "
        f"({time.perf_counter() - t0:.1f} s)")
    with blz.Engine(0) as eng:
        if args.tune:
            eng.set_tuning(**json.loads(args.tune))
        eng.upload_db(res, offs)
        qhs = [eng.prepare(q) for q in queries]
        for i in range(args.warmup):
            got, st = eng.run_batch(qhs, allow_overflow=True, fetch=False)
        eng.set_tuning(**{**eng.get_tuning(), "graph": 1})  # separate stage times
        lat, seeds, gapped, cells = [], [], [], []
        for i in range(args.steps):
            t = time.perf_counter()
            got, st = eng.run_batch(qhs, allow_overflow=True)
            lat.append(time.perf_counter() - t)
            seeds.append(st["ms_seed"])
            gapped.append(st["ms_gapped"])
            cells.append(st["dp_cells"])
            log(f"step {i}: seed {st['ms_seed']:.1f} ms gapped {st['ms_gapped']:.1f} ms, {st['dp_cells']} cells, "
                f"{st['overflow_subjects']} overflow pairs")
        overflow = eng.overflow_pairs()
        tuning = eng.get_tuning()
        for qh in qhs:
            eng.free_query(qh)
    exact, cpu = None, None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle
        threads = min(16, os.cpu_count() or 1)
        sample = [0, nq // 2]
        t = time.perf_counter()
        exact = True
        done_res = 0
        for qi in sample:
            counts, hsps, _ = oracle.prelim_db(queries[qi], res, offs, nthreads=threads)
            want = oracle.as_lists(counts, hsps)
            have = got[qi].as_dict()
            ovf = {s for (q, s) in overflow if q == qi}
            exact &= sorted(set(want) - ovf) == sorted(set(have) - ovf) and all(
                np.array_equal(want[k], have[k]) for k in want if k not in ovf)
            done_res += n_res
        dt = time.perf_counter() - t
        cpu = {"value": done_res / dt, "unit": "residues/s", "cores": threads, "kind": "port",
               "sample": f"queries {sample} of the batch against the whole DB ({done_res} query-residue pairs, "
                         f"{dt:.1f} s), CPU restatement of the NCBI prelim stage (oracle/) on {threads} threads"}
        log(f"cpu {cpu['value']:.3e} residues/s; GPU lists of queries {sample} bit-exact: {exact}")
    step_s = float(np.mean(lat))
    gapped_ms = float(np.mean(gapped))
    out = {
        "metric": METRIC, "value": n_res * nq / step_s, "unit": "residues/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000 * step_s, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{nq} queries of 1,000-2,000 aa vs {n_sub} subjects (log-normal median 800, sigma 0.8, "
                               "clip 30-35,000), 20% homologs at 30-45% identity, 5% of residues in repeat blocks "
                               "of period 3-10; one batched run per step",
                   "queries": nq, "subjects": n_sub, "residues": n_res},
        "gcups": float(np.mean(cells)) / (gapped_ms * 1e-3) / 1e9 if gapped_ms > 0 else None,
        "dp_cells": int(np.mean(cells)), "kernel_ms": {"k_seed": float(np.mean(seeds)), "k_gapped": gapped_ms},
        "overflow_pairs": len(overflow), "bit_exact_sample": exact, "tuning": tuning,
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


TB_SOURCES = ["blz_traceback.hip", "blz_dev_common.h", "blz_device.h", "blz_common.h"]


def tb_source_hash() -> str:
    import hashlib

    h = hashlib.sha256()
    for name in TB_SOURCES:
        h.update((ROOT / "blaze-sc-ad_amd" / "csrc" / name).read_bytes())
    return h.hexdigest()[:16]


def tb_roofline(kernel_ms: float, cells: int, pairs: int) -> dict:
    """k_traceback's roofline: VALU instruction issue, priced with the instruction count of its
    PMC profile on this workload and build (profiles/k_traceback_issue.json), when one matches."""
    peak, basis = valu_peak()
    out = {"kernel": "k_traceback", "bound": "valu", "unit": "G VALU wave-instructions/s", "achieved": None,
           "peak": peak, "peak_basis": basis, "frac": None, "traffic": None, "cells": cells, "pairs": pairs}
    prof_path = ROOT / "profiles" / "k_traceback_issue.json"
    if prof_path.exists():
        prof = json.loads(prof_path.read_text())
        if prof.get("source_sha") == tb_source_hash() and prof.get("pairs") == pairs:
            ach = prof["valu_insts_per_call"] / (kernel_ms * 1e-3) / 1e9
            out.update(achieved=ach, frac=ach / peak, valu_insts_per_call=prof["valu_insts_per_call"],
                       valu_per_cell=prof["valu_insts_per_call"] / max(1, cells), profile=prof.get("source"),
                       traffic=prof.get("bytes_per_call"))
    return out


def run_traceback(args) -> int:
    """Traceback of 256 queries' hit lists: queries log-uniform 50-3,000 aa, each with about 500
    planted homologs (35-70% identity) among 2,000,000 subjects; the prelim batch runs once with
    each query's e-value reap, then per step one blz_traceback call traces every query's hit list
    (expect 10, 500 subjects, no composition-based statistics) with the query's own cutoff and the
    final X-drop (25 bits)."""
    nq = args.queries or 256
    rng = np.random.default_rng(args.seed)
    lens = np.exp(rng.uniform(np.log(50), np.log(3000), nq)).astype(int)
    queries = [blz.synth_query(1000 + i, int(L)) for i, L in enumerate(lens)]
    t0 = time.perf_counter()
    spec = blz.synth_spec(seed=args.seed, n_subjects=args.subjects, len_median=300, len_sigma=0.8, len_min=30,
                          len_max=5000, homolog_frac=min(1.0, 500.0 * nq / args.subjects), ident_min=0.35,
                          ident_max=0.70)
    res, offs, hom = blz.synth_db(spec, queries)
    n_sub, n_res = len(offs) - 1, int(offs[-1])
    log(f"workload built: {nq} queries, {n_sub} subjects, {int((hom >= 0).sum())} homologs "
        f"({time.perf_counter() - t0:.1f} s)")
    opts = blz.stats_options(expect=10.0, composition_based_stats=0)
    stats, params = [], []
    for q in queries:
        p, ss = blz.search_setup(q, n_res, n_sub, opts)
        params.append(p)
        stats.append(ss)
    with blz.Engine(0) as eng:
        if args.tune:
            eng.set_tuning(**json.loads(args.tune))
        eng.upload_db(res, offs)
        qhs = [eng.prepare(q, p) for q, p in zip(queries, params)]
        eng.run_batch(qhs, fetch=False)
        lists = [eng.hitlist(i, ss)[0] for i, ss in enumerate(stats)]
        cutoffs = [s.reap_score for s in stats]
        tbp = blz.tb_params(x_dropoff=stats[0].x_final)
        n_pairs = sum(len(r.oids) for r in lists)
        n_pre = sum(len(r.hsps) for r in lists)
        log(f"hit lists: {n_pairs} pairs, {n_pre} prelim HSPs; final X-drop {tbp.x_dropoff}")
        for i in range(args.warmup):
            got, st = eng.traceback(qhs, lists, tbp, cutoffs)
        calls, kms, wrapper = [], [], []
        for i in range(args.steps):
            t = time.perf_counter()
            got, st = eng.traceback(qhs, lists, tbp, cutoffs)
            wrapper.append(time.perf_counter() - t)
            calls.append(st["ms_total"])
            kms.append(st["ms_main"] + st["ms_wide"])
            log(f"step {i}: {st['ms_total']:.1f} ms call ({1000 * wrapper[-1]:.1f} ms with the Python wrapper), "
                f"{kms[-1]:.2f} ms kernels ({st['ms_main']:.2f} main + {st['ms_wide']:.2f} wide, {st['wide_pairs']} "
                f"wide pairs), {st['hsps']} HSPs, {st['dp_cells']} cells, {st['runs']} runs")
        for qh in qhs:
            eng.free_query(qh)
    call_ms = float(np.median(calls))
    kernel_ms = float(np.median(kms))
    cpu, exact = None, None
    if not args.no_cpu:
        cpu, exact = tb_cpu_baseline(queries, res, offs, lists, got, tbp, cutoffs, args.cpu_budget)
        log(f"cpu baseline {cpu['value']:.1f} pairs/s on {cpu['cores']} threads; sample bit-exact: {exact}")
    out = {
        "metric": "traceback pairs/s (gapped alignment with traceback and edit scripts, bit-exact)",
        "value": n_pairs / (call_ms * 1e-3), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": call_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"hit lists (expect 10, 500 subjects, no composition-based statistics) of {nq} "
                               f"queries, log-uniform 50-3,000 aa (seed {args.seed}), each with ~500 planted homologs "
                               f"at 35-70% identity among {n_sub} subjects; one blz_traceback call per step (pairs "
                               "and prelim HSPs up, edit scripts back on the host)",
                   "queries": nq, "pairs": n_pairs, "prelim_hsps": n_pre, "x_dropoff": tbp.x_dropoff},
        "step": "wall time of the blz_traceback C call (upload, kernels, download, compaction); "
                "python_wrapper_ms adds the ctypes wrapper's per-pair dicts",
        "python_wrapper_ms": 1000 * float(np.median(wrapper)),
        "traced_hsps": st["hsps"], "dp_cells": st["dp_cells"], "runs": st["runs"], "kernel_ms": kernel_ms,
        "gcups": st["dp_cells"] / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else None,
        "wide_pairs": st["wide_pairs"], "bit_exact_sample": exact,
        "roofline": tb_roofline(kernel_ms, st["dp_cells"], n_pairs), "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def tb_cpu_baseline(queries, res, offs, lists, got, params, cutoffs, budget_s: float, runs: int = 5):
    """The oracle's traceback (oracle/traceback.c) on a stride sample of the same pairs, on all the
    host's cores and on one thread (median of `runs` after a warm-up each); the sample's GPU
    results are checked against it."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from concurrent.futures import ThreadPoolExecutor

    import oracle

    host = host_cpu()
    threads = host["nproc"]
    jobs = [(qi, oid, hs) for qi, r in enumerate(lists) for oid, hs in r.as_dict().items()]

    def one(job):
        qi, oid, hs = job
        s = res[int(offs[oid]):int(offs[oid + 1])]
        return oracle.traceback_subject(queries[qi], s, hs, x_final=params.x_dropoff, cutoff=cutoffs[qi])

    def run(sample, th):
        with ThreadPoolExecutor(th) as ex:
            return list(ex.map(one, sample))

    t = time.perf_counter()
    probe = jobs[:: max(1, len(jobs) // 200)]
    run(probe, 1)
    rate1 = len(probe) / max(1e-6, time.perf_counter() - t)
    per_run = budget_s / 2 / (runs + 1)
    s1 = jobs[:: max(1, int(len(jobs) / max(1.0, rate1 * per_run)))]
    run(s1, 1)
    t1, _ = _timed_runs(lambda: run(s1, 1), runs)
    sa = jobs[:: max(1, int(len(jobs) / max(1.0, rate1 * threads * 0.8 * per_run)))]
    run(sa, threads)
    ta, outs = _timed_runs(lambda: run(sa, threads), runs)
    exact = True
    for (qi, oid, _), (wt, wr) in zip(sa, outs):
        gt, gr = got[qi][oid]
        exact &= gt.tolist() == wt.tolist() and gr.tolist() == wr.tolist()
    return {
        "value": len(sa) / ta, "unit": "pairs/s", "cores": threads, "kind": "port", "cpu_model": host["model"],
        "nproc": host["nproc"], "runs": runs,
        "sample": f"every {len(jobs) // max(1, len(sa))}th of the {len(jobs)} pairs ({len(sa)} pairs), median "
                  f"{ta:.2f} s of {runs} runs on {threads} threads; CPU restatement of NCBI's traceback (oracle/)",
        "single_thread": {"value": len(s1) / t1, "unit": "pairs/s", "cores": 1,
                          "sample": f"{len(s1)} pairs, median {t1:.2f} s of {runs} runs"},
    }, exact


def _play(srv, queries, arrivals):
    """Submits queries[i] at arrivals[i] seconds from now (from a client thread) and waits for all;
    returns per-request info dicts and the makespan."""
    import threading

    tickets = [None] * len(queries)
    t0 = time.perf_counter()

    def client():
        for i, a in enumerate(arrivals):
            d = t0 + a - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(queries[i])

    th = threading.Thread(target=client)
    th.start()
    th.join()
    infos = []
    for tk in tickets:
        infos.append(srv.wait(tk))
        srv.release(tk)
    return infos, time.perf_counter() - t0


def run_serve(args) -> int:
    """Poisson arrivals at each rate, served with continuous batching and with one query per
    launch, on the same trace."""
    rates = [float(r) for r in args.rates.split(",")]
    n_subjects = args.subjects if args.subjects != 2_000_000 else 1_000_000
    rng = np.random.default_rng(args.seed)
    spec = blz.synth_spec(seed=args.seed, n_subjects=n_subjects, len_median=300, len_sigma=0.8, len_min=30,
                          len_max=5000, homolog_frac=0.01, ident_min=0.35, ident_max=0.70)
    traces = []
    for k, rate in enumerate(rates):
        n = max(1, int(rate * args.duration))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
        lens = np.clip(np.exp(rng.normal(np.log(350), 0.8, n)), 30, 4000).astype(int)
        traces.append((rate, arrivals, [blz.synth_query(5000 + 1000 * k + i, int(L)) for i, L in enumerate(lens)]))
    t0 = time.perf_counter()
    res, offs, hom = blz.synth_db(spec, traces[0][2][:50])
    n_sub, n_res = len(offs) - 1, int(offs[-1])
    log(f"DB built: {n_sub} subjects, {n_res} residues ({time.perf_counter() - t0:.1f} s)")
    forms = {"batched": dict(slots=args.slots, max_batch=64, max_batch_letters=32768),
             "plain": dict(slots=1, max_batch=1)}
    results = []
    with blz.Engine(0) as eng:
        tune = json.loads(args.tune) if args.tune else None
        eng.upload_db(res, offs)
        for rate, arrivals, queries in traces:
            row = {"rate_qps": rate, "queries": len(queries), "letters": int(sum(len(q) for q in queries))}
            for form, cfg in forms.items():
                with blz.Server(eng, tuning=tune, **cfg) as srv:
                    warm = (queries * 200)[: 3 * 64]  # three full batches: every slot sizes its buffers
                    _play(srv, warm, np.zeros(len(warm)))
                    infos, span = _play(srv, queries, arrivals)
                    st = srv.stats()
                lat = np.array([i["latency_us"] for i in infos]) / 1000.0
                row[form] = {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                             "residues_per_s": len(queries) * n_res / span, "makespan_s": span,
                             "mean_batch": float(np.mean([i["batch_queries"] for i in infos])),
                             "device_ms": st["device_ms"], "failed": int(sum(i["status"] != 0 for i in infos))}
                log(f"rate {rate:g} q/s {form}: p50 {row[form]['p50_ms']:.1f} ms p99 {row[form]['p99_ms']:.1f} ms, "
                    f"{row[form]['residues_per_s']:.3e} residues/s, mean batch {row[form]['mean_batch']:.1f}")
            results.append(row)
    top = results[-1]
    out = {
        "metric": METRIC, "value": top["batched"]["residues_per_s"], "unit": "residues/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": 1000 * top["batched"]["makespan_s"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"Poisson query streams at {args.rates} q/s for {args.duration:g} s each (lengths "
                               f"log-normal median 350, sigma 0.8, clip 30-4,000) vs a resident {n_sub}-subject DB "
                               f"({n_res} residues); serving loop with continuous batching ({args.slots} slots, <= 64 queries) "
                               "against the same loop with one query per launch",
                   "subjects": n_sub, "residues": n_res},
        "latency_ms": {"p50": top["batched"]["p50_ms"], "p99": top["batched"]["p99_ms"]},
        "rates": results, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
