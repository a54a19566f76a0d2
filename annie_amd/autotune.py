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
               for classes of transform lengths 2^9 .. 2^13.

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


def candidates(count):
    slots = sorted({count, max(1, count // 2)})
    launched = [{"graphs": g, "flush_ticks": f, "max_slots": s, "run_steps": 0}
                for g, f, s in itertools.product((False, True), (8, 16, 64), slots)]
    runs = [{"graphs": False, "flush_ticks": 16, "max_slots": s, "run_steps": r, "run_min_active": 1}
            for r, s in itertools.product((8, 32, 128), slots)]
    return launched + runs


def tune_class(tests, seconds=0.5, device=0, configs=None, engine="gpu"):
    """Measure each configuration on `tests` (one length class); returns
    (best config, [(config, square_mul/s), ...])."""
    configs = configs or candidates(len(tests))
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
    best, scores = tune_class(tests, seconds=seconds, device=device, configs=configs, engine=engine)
    cache[k] = {"best": best, "scores": [[c, round(v, 1)] for c, v in scores]}
    save_cache(cache, path)
    return best
