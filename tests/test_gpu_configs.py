"""Whole tests at the sizes of the benchmark configurations, on the GPU path,
against GMP results committed in tests/golden/ (mpz_powm, and for the Proth
k*2^3,000,000+1 test an exact special-form squaring chain):

* the seeded ~332,000-bit Fermat test (plan K=4, LN=13), whose fixture also
  holds the reference's own CPU-path result for the same N;
* the seeded Proth k*2^3,000,000+1 Euler test (plan K=4, pck=2, LN=16,
  lbase=27), run in chunks through checkpoint files and resumed;
* a seeded sample of the 4,096-test batched Proth workload (three tests of
  every length class, n in [2e4, 4e5]), batched with continuous admission;
* a served mix (arrival times, continuous batching) with two requests above
  10^6 bits, and the same mix under an HBM budget that forces queued
  admission.
"""
import json
import os

import pytest

import annie_amd
from tests.helpers import proth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _load(name):
    path = os.path.join(GOLD, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} not generated")
    return json.load(open(path))


def _say(capsys, msg):
    with capsys.disabled():
        print(msg, flush=True)


def _expect(r, ent, what):
    if "factor" in ent:
        assert r["status"] == "factor", (what, r)
        assert r["message"].startswith(ent["factor"]), (what, r)
        return
    assert r["status"] == "success" and r["gl_ok"], (what, r)
    assert r["a"] == ent["a"], (what, r)
    assert (r["res64"], r["is_prp"]) == (ent["res64"], ent["is_prp"]), (what, r, ent)


def test_cfg1_fermat_332000_bits(gpu, capsys):
    fx = _load("config_res64.json")["cfg1"]
    from tests.golden.make_config_fixtures import cfg1_N
    N, cand = cfg1_N()
    assert cand == fx["candidate"] and N.bit_length() == 332000
    assert annie_amd.select_plan(N, 3) == (4, 2, 13, 28)
    _say(capsys, "case cfg1: 332,000-bit Fermat test, base 3")
    r = annie_amd.check(N, a=3, gmp_threshold_digits=1, min_ln=5)
    _say(capsys, f"  {r['message'][-80:]} ({r['seconds']:.1f} s)")
    _expect(r, fx, "cfg1")
    ref = fx.get("reference_check")
    if ref:  # the reference's annie::check on its CPU transform, same N
        assert f"res64 = {fx['res64']:016X}" in ref["results"]


def test_cfg2_proth_3M_prefix_matches_gmp_fixture(gpu, capsys):
    """The plan class of k*2^3,000,000+1 (K=4, pck=2, LN=16, lbase=27) on the
    test's own exponent bits: after bits(k) + j square_mul steps the register
    holds a^(k 2^j) mod N, whose low 64 bits the fixture holds (GMP) for
    j = 1,000, 20,000 and 100,000."""
    fx = _load("cfg2_proth.json")
    k, a = fx["k"], fx["a"]
    N = proth(k, fx["n"])
    pts = sorted((k.bit_length() + p["j"], p["low64"]) for p in fx["prefix"] if p["j"] <= 100_000)
    e = (N - 1) >> 1
    bits = [int(c) for c in format(e >> (e.bit_length() - pts[-1][0]), "b")]
    T = annie_amd.Transform(N, a, min_ln=10)
    try:
        assert not T.init_failed() and (T.K, T.pck, T.LN, T.lbase) == tuple(fx["plan"])
        T.set1()
        done = 0
        for it, low in pts:
            T.square_mul_bits(bits[done:it], a)
            done = it
            got = T.get(0) & ((1 << 64) - 1)
            _say(capsys, f"case cfg2 prefix: iteration {it}, low 64 bits {got:016X}")
            assert got == low, (it, f"{got:016X}", f"{low:016X}")
    finally:
        T.close()


@pytest.mark.long
def test_cfg2_proth_3M_matches_gmp_fixture(gpu, capsys, tmp_path):
    """The Euler test of k*2^3,000,000+1, about 3 million squarings, run in
    chunks of 600,000 loop iterations: each chunk ends with a checkpoint in
    the reference's file format and the next resumes from it."""
    fx = _load("cfg2_proth.json")
    N = proth(fx["k"], fx["n"])
    assert annie_amd.select_plan(N, fx["a"]) == tuple(fx["plan"]) == (4, 2, 16, 27)
    chunks = 0
    while True:
        r = annie_amd.check(N, is_euler=True, workdir=str(tmp_path), quit_after=600_000, min_ln=10)
        chunks += 1
        _say(capsys, f"case cfg2 chunk {chunks}: {r['status']} after {r['iterations']} iterations, "
                     f"{r['seconds']:.1f} s in all (resumed={r['resumed']})")
        if r["status"] != "aborted":
            break
        assert chunks < 10
    assert chunks >= 5 and r["resumed"]
    assert r["status"] == "success" and r["gl_ok"]
    assert r["a"] == fx["a"]
    assert (r["res64"], r["is_prp"]) == (fx["res64"], fx["is_prp"])


def _cfg4_tests(fx):
    return [{"N": proth(e["k"], e["n"]), "is_euler": True} for e in fx]


def _batch_key(plan):
    """The batch driver's batch of a plan: (K, pck, LN, digit-width class), plans
    that differ only in lbase within a width class share a batch."""
    K, pck, LN, lbase = plan
    return (K, pck, LN, (pck * lbase) >> 5)


def test_cfg4_sample_batched_matches_gmp(gpu, capsys):
    fx = _load("config_res64.json")["cfg4"]
    plans = {tuple(e["plan"]) for e in fx}
    assert {p[2] for p in plans} == {9, 10, 11, 12, 13}
    tests = _cfg4_tests(fx)
    _say(capsys, f"case cfg4 sample: {len(tests)} Proth tests in {len(plans)} length classes")
    res, st = annie_amd.batch_run(tests, engine="gpu")
    _say(capsys, f"  {st}")
    assert st["rc"] == 0 and st["classes"] == len({_batch_key(p) for p in plans})
    for e, r in zip(fx, res):
        _expect(r, e, f"{e['k']}*2^{e['n']}+1")


def _cfg5_tests():
    from tests.golden.make_config_fixtures import cfg5_mix
    mix = cfg5_mix()
    for i, t in enumerate(mix):
        t["arrival"] = 0.05 * i
    return mix


def test_cfg5_served_mix_matches_gmp(gpu, capsys):
    fx = _load("config_res64.json")["cfg5"]["requests"]
    mix = _cfg5_tests()
    assert len(mix) == len(fx) and max(t["N"].bit_length() for t in mix) > 1_000_000
    _say(capsys, f"case cfg5 served mix: {len(mix)} requests, largest {max(e['bits'] for e in fx)} bits")
    res, st = annie_amd.batch_run(mix, engine="gpu", realtime=True)
    _say(capsys, f"  {st}")
    for e, r, t in zip(fx, res, mix):
        _expect(r, e, f"request {e['index']} ({e['bits']} bits)")
        assert r["admitted"] >= t["arrival"] - 1e-6


def test_hbm_budget_forces_queued_admission(gpu, capsys):
    """A budget of a few MiB leaves one slot per length class: tests of a
    class wait for the one before them, and every result is unchanged."""
    fx = _load("config_res64.json")["cfg4"]
    ents = [e for e in fx if e["plan"][2] <= 10][:9]
    tests = _cfg4_tests(ents)
    classes = {_batch_key(e["plan"]) for e in ents}
    per_class = {}
    for e in ents:
        per_class[_batch_key(e["plan"])] = per_class.get(_batch_key(e["plan"]), 0) + 1
    assert max(per_class.values()) >= 2
    # slots of a class = max(1, floor((budget / classes - fixed) / per_slot)):
    # below classes * (fixed + 2 per_slot) for every class, that is one slot
    sizes = [annie_amd.batch_bytes(K, LN) for K, pck, LN, width in classes]
    budget_mb = len(classes) * min(fixed + 1.5 * slot for slot, fixed in sizes) / 2 ** 20
    res, st = annie_amd.batch_run(tests, engine="gpu", hbm_budget_mb=budget_mb)
    _say(capsys, f"case budget {budget_mb:.1f} MiB: {st}")
    assert st["slots"] == len(classes)
    for e, r in zip(ents, res):
        _expect(r, e, f"{e['k']}*2^{e['n']}+1")
    # a test of a class that holds one slot is admitted only after the
    # previous test of that class finished
    order = {}
    for e, r in zip(ents, res):
        if r["status"] != "factor":
            order.setdefault(_batch_key(e["plan"]), []).append(r)
    for rs in order.values():
        rs = sorted(rs, key=lambda x: x["admitted"])
        for prev, nxt in zip(rs, rs[1:]):
            assert nxt["admitted"] >= prev["finished"] - 1e-6
