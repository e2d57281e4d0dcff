"""Autotuner for graph-captured plans.

A plan's schedule has two knobs: how many internal streams it spreads work
over and how many instance chunks each task is cut into.  Chunks let the
latency-bound hash chains of one chunk (few waves) run beside the
throughput-bound products of another; too many chunks add launches and
shrink the product kernels.  Kernel launch-shape knobs of the engine
(``Engine.set_tuning``, e.g. the Saber product's instances per workgroup)
are tuned after them, one knob at a time at the best schedule (coordinate
descent: ``refine`` proposes the extra candidates).  The best point depends
on scheme, batch size and op mix, so it is measured: every candidate plan is
built, replayed and timed with HIP events on the engine's stream; the winner
is cached in a JSON file keyed by device name and workload signature.

The cache lives at $DPKEM_TUNE_CACHE (default ~/.cache/dpkem/tune.json).
"""
from __future__ import annotations

import json
import os
import statistics
import time

DEFAULT_CANDIDATES = ((1, 1), (2, 1), (4, 1), (8, 1), (2, 2), (4, 2), (4, 4), (8, 4), (4, 8), (8, 8))


def cache_path() -> str:
    return os.environ.get("DPKEM_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "dpkem", "tune.json"))


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only disables caching


def time_plan(torch, eng, plan, reps: int) -> float:
    """Median ms of `reps` graph replays on the engine's (torch's current) stream."""
    stream = torch.cuda.current_stream()
    plan.launch()  # first replay warms caches and code objects
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        plan.launch()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    return statistics.median(times)


def tune(torch, eng, key: str, build_plan, candidates=DEFAULT_CANDIDATES, reps: int = 3,
         use_cache: bool = True, budget_s: float = 60.0, refine=None):
    """Pick the candidate (a tuple of knob values, e.g. (nstreams, chunks)) for
    the workload that `build_plan(*candidate)` builds.

    `refine(best)`, if given, returns further candidates that are timed after
    the first sweep (within the same budget).  Returns (best, {candidate: ms}).
    `key` names the workload (scheme, op mix, batch size, knob set); the
    device name is added to it.
    """
    dev = torch.cuda.get_device_name(0)
    full_key = f"{dev}|{key}"
    db = _load() if use_cache else {}
    if use_cache and full_key in db:
        e = db[full_key]
        return tuple(e["best"]), {tuple(map(int, k.split("x"))): v for k, v in e["ms"].items()}
    results = {}
    t_end = time.perf_counter() + budget_s

    def sweep(cands):
        for cand in cands:
            if cand in results:
                continue
            if results and time.perf_counter() > t_end:
                break
            plan = build_plan(*cand)
            try:
                results[cand] = time_plan(torch, eng, plan, reps)
            finally:
                plan.close()

    sweep([tuple(c) for c in candidates])
    if refine is not None:
        sweep([tuple(c) for c in refine(min(results, key=results.get))])
    best = min(results, key=results.get)
    db = _load()
    db[full_key] = {"best": list(best), "ms": {"x".join(map(str, c)): round(v, 4) for c, v in results.items()}}
    if use_cache:
        _store(db)
    return best, results
