"""Writes tests/golden/cfg3_prefix.json: for the ~40M-digit benchmark N
(bench.workload_N), the canonical residue 3^(top t bits of N-1) mod N after
t loop iterations, computed with GMP's mpz_powm (the textbook oracle), for a
few t.  Stored as res64 and SHA-256 of the little-endian bytes of the residue.
Run once in the build container (GMP on the CPU takes minutes)."""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import oracle  # noqa: E402


def main():
    iters = [int(x) for x in sys.argv[1:]] or [100, 1000]
    N = bench.workload_N()
    e = N - 1
    out = {"bits": bench.BITS, "seed": bench.SEED, "base": bench.BASE, "candidate": 0, "entries": []}
    for t in iters:
        t0 = time.time()
        r = oracle.powm(bench.BASE, e >> (e.bit_length() - t), N)
        nbytes = (N.bit_length() + 7) // 8
        out["entries"].append({"iterations": t, "res64": r & ((1 << 64) - 1),
                               "sha256": hashlib.sha256(r.to_bytes(nbytes, "little")).hexdigest(),
                               "seconds": round(time.time() - t0, 1)})
        print(out["entries"][-1], flush=True)
    with open(os.path.join(HERE, "cfg3_prefix.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
