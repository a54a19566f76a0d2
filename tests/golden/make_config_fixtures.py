"""Writes tests/golden/config_res64.json: GMP results of whole tests at the
sizes of the benchmark configurations, for the GPU parity tests.

* cfg1: the seeded random odd ~332,000-bit N (random.Random(1), top bit and
  bit 0 set, first candidate that initialises under both the GPU plan and the
  reference's CPU plan), Fermat base 3.  res64 = low 64 bits of 3^(N-1) mod N
  (mpz_powm).  When oracle/_ref is built, the reference's own annie::check
  (CPU transform, src/annie.h:217-430) is run on the same N in a scratch
  directory; its res64 line and wall time are stored beside (ref_*).
* cfg4: a seeded sample of the 4,096-test Proth workload (k odd uniform in
  [3, 2^20), n log-uniform in [2e4, 4e5], seed 4): the first three tests of
  every transform-length class, in workload order.
* cfg5: a served mix of Fermat/Proth requests drawn like the serving
  benchmark's (seed 55, bits log-uniform in [1e4, 1.2e5]) plus one Proth and
  one Fermat request above 10^6 bits.

For every test: base a (src/annie.h:249-254), the plan (K, pck, LN, lbase),
whether gcd(N, R) > 1 (the transform then reports a factor), and otherwise
is_prp and res64 from mpz_powm: Fermat x = a^(N-1); Euler x = a^((N-1)/2),
is_prp <=> x = N - 1, res64 from x^2 (src/annie.h:343-366).
Run once in the build container (minutes, a process pool)."""
import json
import math
import multiprocessing as mp
import os
import random
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import oracle  # noqa: E402

MASK = (1 << 64) - 1


def R_of(N, a):
    K, pck, LN, lbase = oracle.plan(N, a)
    return (1 << ((pck * lbase) << LN)) - 1, (K, pck, LN, lbase)


def cfg1_N():
    """Seeded 332,000-bit N whose GPU transform (plan K=4, LN=13) and the
    reference CPU transform (K=5, pck=4, LN=13, lbase=17) both initialise."""
    rng = random.Random(1)
    cand = 0
    while True:
        N = rng.getrandbits(332000) | (1 << 331999) | 1
        R, _ = R_of(N, 3)
        Rc = (1 << (4 * 17 << 13)) - 1
        if math.gcd(N, R) == 1 and math.gcd(N, Rc) == 1:
            return N, cand
        cand += 1


def cfg4_sample(per_class=3):
    tests = bench.proth_batch_workload(4096, 4)
    seen, out = {}, []
    for i, t in enumerate(tests):
        a = bench.euler_base(t["N"])
        key = oracle.plan(t["N"], a)
        if seen.get(key, 0) < per_class:
            seen[key] = seen.get(key, 0) + 1
            out.append({"index": i, "k": t["k"], "n": t["n"]})
    return out


def cfg5_mix():
    rng = random.Random(55)
    reqs = []
    for _ in range(10):
        r = bench.serve_request(rng, 1e4, 1.2e5)
        reqs.append(r)
    k = rng.randrange(3, 1 << 32) | 1
    reqs.append({"N": k * (1 << (1_050_000 - k.bit_length())) + 1, "is_euler": True})
    reqs.append({"N": rng.getrandbits(1_010_000) | (1 << 1_009_999) | 1, "a": 3})
    return reqs


def solve(job):
    """(tag, index, N, is_euler) -> fixture entry."""
    tag, idx, N, is_euler = job
    t0 = time.time()
    a = bench.euler_base(N) if is_euler else 3
    R, plan = R_of(N, a)
    ent = {"tag": tag, "index": idx, "bits": N.bit_length(), "is_euler": is_euler, "a": a, "plan": list(plan)}
    if math.gcd(N, R) > 1:
        ent["factor"] = str(math.gcd(N, R))
    else:
        if is_euler:
            x = oracle.powm(a, (N - 1) >> 1, N)
            ent["is_prp"] = x == N - 1
            ent["res64"] = (x * x % N) & MASK
        else:
            x = oracle.powm(a, N - 1, N)
            ent["is_prp"] = x == 1
            ent["res64"] = x & MASK
    ent["seconds"] = round(time.time() - t0, 1)
    print(tag, idx, ent["bits"], ent.get("res64", ent.get("factor")), ent["seconds"], flush=True)
    return ent


def ref_cfg1(N):
    """The reference's annie::check (CPU transform) on N: res64 line, seconds."""
    if not oracle.ref_available():
        return None
    arr, n = oracle.limbs(N)
    with tempfile.TemporaryDirectory() as d:
        t0 = time.time()
        rc = oracle.ref().ref_check(arr, n, 0, 3, d.encode())
        dt = time.time() - t0
        text = open(os.path.join(d, "results.txt")).read()
    return {"rc": rc, "results": text.strip().splitlines()[-1], "seconds": round(dt, 1)}


def main():
    procs = int(sys.argv[1]) if len(sys.argv) > 1 else 2
    N1, cand1 = cfg1_N()
    sample = cfg4_sample()
    mix = cfg5_mix()
    jobs = [("cfg1", 0, N1, False)]
    jobs += [("cfg4", s["index"], s["k"] * (1 << s["n"]) + 1, True) for s in sample]
    jobs += [("cfg5", i, r["N"], bool(r.get("is_euler"))) for i, r in enumerate(mix)]
    with mp.Pool(procs) as pool:
        ents = pool.map(solve, jobs, chunksize=1)
    out = {"cfg1": {"seed": 1, "bits": 332000, "candidate": cand1, **ents[0]},
           "cfg4": [dict(e, k=s["k"], n=s["n"]) for e, s in zip(ents[1:1 + len(sample)], sample)],
           "cfg5": {"seed": 55, "requests": ents[1 + len(sample):]}}
    out["cfg1"]["reference_check"] = ref_cfg1(N1)
    print("cfg1 reference:", out["cfg1"]["reference_check"], flush=True)
    with open(os.path.join(HERE, "config_res64.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
