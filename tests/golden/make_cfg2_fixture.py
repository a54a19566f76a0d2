"""Writes tests/golden/cfg2_proth.json: the Euler (Proth) test of
N = k*2^3,000,000 + 1 for the seeded 32-bit odd k (random.Random(2),
getrandbits(32) | 1 | 2^31, first candidate whose transform initialises,
i.e. gcd(N, R) = 1 for R = 2^(pck*lbase*2^LN) - 1 of the reference's plan).

The residues come from proth_res64.cpp (GMP, exact integer arithmetic: a^k
by mpz_powm, then modular squarings with a special-form reduction checked
against mpz_mod).  Stored: k, n, base a, is_prp, res64 (low 64 bits of
a^(N-1) mod N, src/annie.h:355-366) and the low 64 bits of a^(k 2^j) mod N at
a few j.  Run once in the build container; takes hours (six threads).
Usage: make_cfg2_fixture.py [STATE_FILE] (resumable; default under the
system temporary directory)."""
import json
import math
import os
import random
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

N_EXP = 3_000_000
REPORT = [1000, 20000, 100000, 1000000]


def seeded_k():
    rng = random.Random(2)
    while True:
        k = rng.getrandbits(32) | 1 | (1 << 31)
        N = k * (1 << N_EXP) + 1
        a = oracle.base(N, True)
        K, pck, LN, lbase = oracle.plan(N, a)
        m = pck * lbase << LN
        r = oracle.powm(2, m, N)
        if math.gcd(N, (r - 1) % N) == 1:
            return k, a, (K, pck, LN, lbase)


def main():
    # Scratch (the generator binary and its resumable state) outside the tree.
    state = sys.argv[1] if len(sys.argv) > 1 else os.path.join(tempfile.gettempdir(), "annie_cfg2", "state.bin")
    os.makedirs(os.path.dirname(state), exist_ok=True)
    k, a, plan = seeded_k()
    print(f"k = {k}, base {a}, plan {plan}", flush=True)
    exe = os.path.join(os.path.dirname(state), "proth_res64")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I/opt/conda/include",
                    os.path.join(HERE, "proth_res64.cpp"), "-o", exe,
                    "-L/usr/lib/x86_64-linux-gnu", "-l:libgmp.so.10"], check=True)
    t0 = time.time()
    out = subprocess.run([exe, str(k), str(N_EXP), state] + [str(j) for j in REPORT],
                         check=True, capture_output=True, text=True).stdout
    res = {"k": k, "n": N_EXP, "plan": list(plan), "seconds": round(time.time() - t0)}
    prefix = []
    for ln in out.splitlines():
        f = ln.split()
        if f[0] == "a":
            res["a"] = int(f[1])
        elif f[0] == "is_prp":
            res["is_prp"] = bool(int(f[1]))
        elif f[0] == "res64":
            res["res64"] = int(f[1], 16)
        elif f[0] == "iter":
            prefix.append({"j": int(f[1]), "low64": int(f[3], 16)})
    if res.get("a") != a or "res64" not in res:
        raise SystemExit(f"unexpected generator output: {out[-500:]}")
    res["prefix"] = prefix
    with open(os.path.join(HERE, "cfg2_proth.json"), "w") as fo:
        json.dump(res, fo, indent=1)
    print(res, flush=True)


if __name__ == "__main__":
    main()
