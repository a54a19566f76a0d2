"""Batched / served test driver (annie_batch_run) on the CPU: the per-test
protocol state machine and the continuous-admission driver run on exact GMP
transforms and must reproduce the oracle's protocol (a restatement of
annie::check, src/annie.h:217-430) test by test, whatever the slot count,
arrival order or mix of classes."""
import random

import pytest

import annie_amd
import oracle
from tests.helpers import proth


def _mix(seed, count):
    rng = random.Random(seed)
    tests = []
    for _ in range(count):
        if rng.random() < 0.6:
            k = rng.randrange(3, 1 << 12) | 1
            tests.append({"N": proth(k, rng.randrange(1200, 3000)), "is_euler": True})
        else:
            tests.append({"N": rng.getrandbits(rng.randrange(2500, 6000)) | 1, "a": 3})
    return tests


def _expect(t):
    return oracle.check(t["N"], is_euler=t.get("is_euler", False), a_ref=t.get("a", 0))


def _assert_matches(tests, res):
    for t, r in zip(tests, res):
        o = _expect(t)
        if o.get("init_failed"):
            assert r["status"] == "factor", r
            continue
        assert r["status"] == "success", r
        assert r["gl_ok"]
        assert r["res64"] == o["res64"] and r["is_prp"] == o["is_prp"] and r["a"] == o["a"], (r, o)


@pytest.mark.parametrize("max_slots", [0, 1, 3])
def test_batch_matches_oracle(max_slots):
    tests = _mix(11, 24)
    res, st = annie_amd.batch_run(tests, engine="gmp", max_slots=max_slots)
    assert st["rc"] == 0
    _assert_matches(tests, res)
    # every square_mul of every test is accounted for
    assert st["squarings"] >= sum(r["iterations"] for r in res)


@pytest.mark.parametrize("run_steps,min_active", [(1, 1), (8, 1), (64, 3)])
def test_squaring_runs_match_oracle(run_steps, min_active):
    """Run mode (up to run_steps squarings per test per tick, ending at a
    Gerbicz-Li snapshot) issues the same square_mul sequence per test."""
    tests = _mix(12, 16)
    base, st0 = annie_amd.batch_run(tests, engine="gmp", max_slots=5)
    res, st = annie_amd.batch_run(tests, engine="gmp", max_slots=5, run_steps=run_steps, run_min_active=min_active)
    assert st["rc"] == 0
    _assert_matches(tests, res)
    assert st["squarings"] == st0["squarings"]
    for a, b in zip(base, res):
        assert (a["iterations"], a["res64"], a["is_prp"]) == (b["iterations"], b["res64"], b["is_prp"])


def test_known_proth_primes_in_a_batch():
    # 3*2^n+1 is prime for n = 2208, 2816, 3168 (OEIS A002253); 3*2^2209+1 is not.
    tests = [{"N": proth(3, n), "is_euler": True} for n in (2208, 2816, 3168, 2209)]
    res, _ = annie_amd.batch_run(tests, engine="gmp", max_slots=2)
    assert [r["is_prp"] for r in res] == [True, True, True, False]
    assert all(r["gl_ok"] for r in res)


def test_served_arrivals_respected():
    tests = _mix(12, 10)
    for i, t in enumerate(tests):
        t["arrival"] = 0.02 * i
    res, _ = annie_amd.batch_run(tests, engine="gmp", max_slots=2, realtime=True)
    _assert_matches(tests, res)
    for r in res:
        if r["status"] == "success":
            assert r["admitted"] >= r["arrival"] - 1e-6
            assert r["finished"] >= r["admitted"]


def test_decided_without_the_loop():
    res, _ = annie_amd.batch_run([{"N": 1000003}, {"N": 10 ** 20 + 2}, {"N": 999999999999989}], engine="gmp")
    assert [r["status"] for r in res] == ["decided"] * 3
    assert res[0]["is_prp"] and not res[1]["is_prp"] and res[2]["is_prp"]


def test_time_limit_aborts_unfinished():
    tests = [{"N": proth(3, 20000) + 0, "is_euler": True}]
    res, st = annie_amd.batch_run(tests, engine="gmp", time_limit=0.2)
    assert res[0]["status"] == "aborted"
    assert 0 < res[0]["iterations"] < 20000


def test_autotuner_picks_a_measured_config(tmp_path):
    from annie_amd import autotune
    tests = _mix(13, 4)
    cfgs = [{"graphs": False, "flush_ticks": f, "max_slots": 0} for f in (4, 16)]
    best, scores = autotune.tune_class(tests, seconds=0.05, configs=cfgs, engine="gmp")
    assert best in cfgs and len(scores) == 2
    path = str(tmp_path / "tune.json")
    b1 = autotune.tuned((4, 2, 10, 28), tests, path=path, seconds=0.05, configs=cfgs, engine="gmp")
    assert autotune.load_cache(path)[autotune.key_of((4, 2, 10, 28))]["best"] == b1
    # cached: no re-measurement needed
    assert autotune.tuned((4, 2, 10, 28), tests, path=path, configs=[{"bogus": 1}], engine="gmp") == b1


def test_hbm_budget_bounds_the_slots_of_each_class():
    """Admission under an HBM budget: each class gets slots whose device
    memory (the engine's own bytes per slot and per batch) fits its equal
    share of the budget, and the tests beyond those slots queue and still
    finish with the oracle's results."""
    tests = [{"N": proth(k, 2000), "is_euler": True} for k in (3, 5, 9, 15, 17, 21, 27, 33)]
    plans = {annie_amd.select_plan(t["N"], oracle.base(t["N"], True)) for t in tests}
    assert len(plans) == 1
    K, pck, LN, lbase = plans.pop()
    per_slot, fixed = annie_amd.batch_bytes(K, LN)
    # NREG registers of K n words plus 2 x 3 digit planes per slot, 4 twiddle tables per batch
    assert per_slot >= (13 * K + 6) * (4 << LN) and fixed == 16 * K << LN
    budget = fixed + 3 * per_slot + per_slot // 2          # room for exactly 3 slots
    res, st = annie_amd.batch_run(tests, engine="gmp", hbm_budget_mb=budget / 1048576.0)
    assert st["slots"] == 3
    assert st["slots"] * per_slot + fixed <= budget
    _assert_matches(tests, res)
    admitted = sorted(r["admitted"] for r in res if r["status"] == "success")
    assert admitted[-1] > admitted[2]  # later tests waited for a free slot


@pytest.mark.parametrize("host_threads", [1, 3, 16])
def test_host_workers_give_the_same_results(host_threads):
    """Classes driven by 1, 3 or 16 host worker threads: the same results
    test by test, every square_mul accounted for."""
    tests = _mix(13, 20)
    base, st0 = annie_amd.batch_run(tests, engine="gmp", max_slots=4, host_threads=1)
    res, st = annie_amd.batch_run(tests, engine="gmp", max_slots=4, host_threads=host_threads)
    assert st["rc"] == 0 and st["squarings"] == st0["squarings"]
    for a, b in zip(base, res):
        assert (a["status"], a["res64"], a["is_prp"], a["gl_ok"], a["iterations"]) == \
            (b["status"], b["res64"], b["is_prp"], b["gl_ok"], b["iterations"])
    _assert_matches(tests, res)
