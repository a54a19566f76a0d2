"""Autotuner for the batched path: per length class, short measured trials of
the batch driver's schedule knobs, the best kept in a JSON cache.

Knobs (all leave results unchanged; only the schedule differs):
  graphs       batched squarings replayed from captured HIP graphs, or issued
               launch by launch;
  flush_ticks  period (in ticks) at which the deferred Gerbicz-Li products of
               all tests of the class are taken together;
  max_slots    tests resident at once (all of the class, or a fraction);
  run_steps    squarings per test per launch in squaring runs (one workgroup
               per test runs whole squarings; 0 = one squaring per launch),
               for classes of transform lengths 2^9 .. 2^13;
  row_lines,   the NTT launch schedule: lines per row-pass workgroup (rows of
  col_cols     up to 2^9 elements) and columns per column-pass workgroup
               (columns of up to 2^9), every pair that fits the length
               (annie_amd.schedules), for the two-pass lengths 2^14 and up.

tune_plain does the same for one test on the plain engine: each fitting
schedule timed on the test's own squarings.

A trial runs the class's own tests for `seconds` of square_mul work (batch
set-up excluded) and scores square_mul operations per second.  The cache key
is the GPU name and the plan (K, pck, LN, lbase).
"""
import itertools
import json
import os

import annie_amd

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "annie_amd", "batch_tune.json")


def _device_name():
    # The engine is built for gfx950 only.
    return "gfx950"


def candidates(count, LN=None):
    slots = sorted({count, max(1, count // 2)})
    launched = [{"graphs": g, "flush_ticks": f, "max_slots": s, "run_steps": 0}
                for g, f, s in itertools.product((False, True), (8, 16, 64), slots)]
    runs = [{"graphs": False, "flush_ticks": 16, "max_slots": s, "run_steps": r, "run_min_active": 1}
            for r, s in itertools.product((8, 32, 128), slots)]
    sched = []
    if LN is not None:
        sched = [{"graphs": False, "flush_ticks": 16, "max_slots": count, "run_steps": 0,
                  "row_lines": rl, "col_cols": cc} for rl, cc in annie_amd.schedules(LN)[1:]]
    return launched + runs + sched


def tune_class(tests, seconds=0.5, device=0, configs=None, engine="gpu", LN=None):
    """Measure each configuration on `tests` (one length class of transform
    length 2^LN); returns (best config, [(config, square_mul/s), ...])."""
    configs = configs or candidates(len(tests), LN)
    scores = []
    for cfg in configs:
        _, st = annie_amd.batch_run(tests, engine=engine, device=device, time_limit=seconds, **cfg)
        busy = st["seconds"] - st["setup_seconds"]
        scores.append((cfg, st["squarings"] / busy if busy > 0 else 0.0))
    best = max(scores, key=lambda cs: cs[1])[0]
    return best, scores


def load_cache(path=DEFAULT_CACHE):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=DEFAULT_CACHE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def key_of(plan, device_name=None):
    return f"{device_name or _device_name()}|K={plan[0]},pck={plan[1]},LN={plan[2]},lbase={plan[3]}"


def tuned(plan, tests, path=DEFAULT_CACHE, seconds=0.5, device=0, retune=False, configs=None, engine="gpu"):
    """Best configuration for `plan`, from the cache or measured now."""
    cache = load_cache(path)
    k = key_of(plan)
    if k in cache and not retune:
        return cache[k]["best"]
    best, scores = tune_class(tests, seconds=seconds, device=device, configs=configs, engine=engine, LN=plan[2])
    cache[k] = {"best": best, "scores": [[c, round(v, 1)] for c, v in scores]}
    save_cache(cache, path)
    return best


def tune_plain(N, a, squarings=400, device=0, min_ln=10):
    """One test on the plain engine: every launch schedule that fits its
    transform length timed over `squarings` square_mul steps with the top
    exponent bits of N - 1 (after a short warm-up of each); returns (best
    (row_lines, col_cols), [((row_lines, col_cols), ms per squaring), ...]).
    The default schedule is (0, 0)."""
    T = annie_amd.Transform(N, a, device=device, min_ln=min_ln)
    try:
        e = N - 1
        nb = min(e.bit_length(), squarings + 32)
        bits = [int(c) for c in format(e >> (e.bit_length() - nb), "b")]
        T.set1()
        scores = []
        for sch in annie_amd.schedules(T.LN):
            T.set_schedule(*sch)
            T.square_mul_bits(bits[:32], a)
            ms = T.time_square_mul_bits(bits[32:], a) / max(1, len(bits) - 32)
            scores.append((sch, ms))
        T.set_schedule(0, 0)
    finally:
        T.close()
    return min(scores, key=lambda sm: sm[1])[0], scores

