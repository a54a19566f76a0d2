#!/usr/bin/env python3
"""Benchmark of the MMGW modular-squaring hot path on one MI355X.

Workload (the ~40M-digit configuration): Fermat PRP test of a seeded random
odd 132,900,000-bit N (top bit set), base 3, with the transform the reference
selects for it (K = 5 primes, n = 2^21, 16-bit digits packed by 4).  A step is
one `square_mul` of the reference's hot loop (src/annie.h:319-341): the
exponent bits are the top bits of e = N - 1, exactly as the test consumes
them.  Registers, twiddles and constants are resident in HBM before timing.

Prints ONE JSON line.  `value` is ms per squaring (lower is better), timed
with HIP events on the engine's stream over `--steps` squarings after
`--warmup` untimed ones.  `roofline` is computed live for the kernel that
takes the most time per squaring (HIP event pairs around its launches over a
profile window of the same workload); `cpu_baseline` is the reference's own
CPU transform (src/transformCPU.h, built from the reference sources into
oracle/_ref) timed on a bounded sample on one host core, run concurrently
while the GPU context initialises.
"""
import argparse
import json
import math
import os
import random
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BITS = 132_900_000
BASE = 3
SEED = 3
HBM_PEAK_GBS = 8000.0
# Product of the small prime factors of R = 2^(2^27) - 1 (Fermat-number
# factors F0..F5): an N sharing one of them cannot be tested with this R.
SMALL_R_FACTORS = 3 * 5 * 17 * 257 * 65537 * 641 * 6700417


def workload_N(bits=BITS, seed=SEED, skip=0):
    """Seeded N: candidate `skip` of the seeded stream that is coprime to the
    small factors of R (the engine re-checks gcd(N, R) exactly)."""
    rng = random.Random(seed)
    idx = -1
    while True:
        N = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        if math.gcd(N, SMALL_R_FACTORS) == 1:
            idx += 1
            if idx == skip:
                return N


def exponent_bits(N, count):
    """The first `count` bits (from the top) of e = N - 1."""
    e = N - 1
    top = e >> (e.bit_length() - count)
    return [int(c) for c in format(top, "b")]


def alg_bytes(kind, U):
    """Algorithmic HBM bytes of one launch, U = K * n * 4 (one register).

    Every NTT pass reads and writes each job's register once (2U per job);
    the fused row pass of a product also reads the transformed operand (U);
    the carry reads the convolution residues (U) and writes a register (U).
    """
    name, jobs = kind.rsplit("x", 1)
    jobs = int(jobs)
    if name.startswith("k_row"):
        mode = name.split(",")[1].rstrip(">")
        return jobs * (3 * U if mode in ("mulc", "mulv") else 2 * U)
    return jobs * 2 * U


PMC_SUMMARY = os.path.join(ROOT, "profiles", "square_mul_40M_digits_pmc.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` and its VALU utilisation from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (counters scaled by a plain copy of known size), or (None, None)."""
    try:
        lines = open(PMC_SUMMARY).read().splitlines()
    except OSError:
        return None, None
    for ln in lines:
        f = ln.split()
        if len(f) < 2 or f[0] != "traffic" or f[1] != kernel:
            continue
        kv = dict(x.split("=", 1) for x in f[2:])
        b = (float(kv["fetch_kib"]) * float(kv["fetch_factor"]) + float(kv["write_kib"]) * float(kv["write_factor"])) * 1024
        return int(round(b)), float(kv["valu_util"]) if "valu_util" in kv else None
    return None, None


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline_worker(skip, seconds):
    """Runs in a child process pinned to one core: the reference transformCPU."""
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    import oracle  # the checker / baseline; never the measured path
    N = workload_N(skip=skip)
    bits = exponent_bits(N, 64)
    kind = "reference" if oracle.ref_available() else "port"
    t0 = time.time()
    T = oracle.RefTransform(N, BASE, 1) if kind == "reference" else oracle.NttTransform(N, BASE)
    init_s = time.time() - t0
    if T.init_failed():
        print(json.dumps({"error": "init failed"}), flush=True)
        return
    T.set1()
    done, t0 = 0, time.time()
    while done < len(bits) and (done < 2 or time.time() - t0 < seconds):
        T.square_mul(BASE if bits[done] else 1)
        done += 1
    dt = time.time() - t0
    info = T.info() if kind == "reference" else "oracle RNS-NTT port"
    print(json.dumps({"ms_per_squaring": dt * 1e3 / done, "squarings": done, "seconds": dt,
                      "init_s": init_s, "kind": kind, "transform": info}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-worker", type=int, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_baseline_worker is not None:
        cpu_baseline_worker(args.cpu_baseline_worker, args.cpu_seconds)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the path runs one test per GPU and this build has no "
              f"multi-GPU mode (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import annie_amd

    # Pick the first seeded N the engine accepts (gcd(N, R) = 1).
    skip = 0
    while True:
        N = workload_N(skip=skip)
        progress(f"candidate {skip}: {N.bit_length()}-bit N; creating the context (N^-1 mod R takes about a minute)")
        t0 = time.time()
        cpu = None
        if rank == 0 and not args.no_cpu_baseline:
            cpu = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", str(skip),
                                    "--cpu-seconds", str(args.cpu_seconds)],
                                   stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        T = annie_amd.Transform(N, BASE)
        init_s = time.time() - t0
        progress(f"context ready in {init_s:.1f} s (init_failed={T.init_failed()})")
        if not T.init_failed():
            break
        if cpu is not None:
            cpu.kill()
            cpu.wait()
        T.close()
        skip += 1

    U = T.K * T.n * 4
    total = args.warmup + args.steps + args.profile_steps
    bits = exponent_bits(N, total)
    T.set1()
    progress(f"warmup: {args.warmup} squarings")
    T.square_mul_bits(bits[:args.warmup], BASE)
    T.sync()
    progress(f"timed region: {args.steps} squarings")
    w0 = time.time()
    ms = T.time_square_mul_bits(bits[args.warmup:args.warmup + args.steps], BASE)
    wall = time.time() - w0
    ms_per_sq = ms / args.steps

    progress(f"{ms_per_sq:.4f} ms/squaring; profile window: {args.profile_steps} squarings")
    T.profile_begin()
    T.square_mul_bits(bits[args.warmup + args.steps:], BASE)
    prof = T.profile_end()
    per_sq = {k: tot / args.profile_steps for k, (cnt, tot) in prof.items()}
    dom = max(prof, key=lambda k: prof[k][1])
    cnt, tot = prof[dom]
    avg_ms = tot / cnt
    achieved = alg_bytes(dom, U) / (avg_ms * 1e-3) / 1e9
    b_alg_sq = 27 * U  # (16P - 5) U with P = 2 two-pass directions at this length
    traffic, valu_util = measured_traffic(dom)

    # A plain register copy of known size (U bytes read, U written) for
    # calibrating HBM byte counters in profiling runs.
    T.copy(1, 0)
    T.sync()

    cpu_res = None
    if cpu is not None:
        progress("waiting for the CPU baseline (reference transformCPU on one core)")
        while cpu.poll() is None:
            time.sleep(10)
            progress("CPU baseline still running")
        out, err = cpu.communicate(timeout=900)
        try:
            cpu_res = json.loads(out.strip().splitlines()[-1])
        except (ValueError, IndexError):
            cpu_res = {"error": (err or out)[-300:]}

    line = {
        "metric": "ms/squaring at ~40M digits",
        "value": round(ms_per_sq, 5),
        "unit": "ms/squaring",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_sq, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "Fermat PRP square_mul loop, seeded random odd 132,900,000-bit N (~40M digits), base 3",
            "bits": BITS, "seed": SEED, "candidate": skip, "K": T.K, "n": T.n, "LN": T.LN, "lbase": T.lbase,
            "pck": T.pck, "parallelism": "single GPU",
        },
        "squarings_per_s": round(1e3 / ms_per_sq, 1),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 5),
        "init_s": round(init_s, 1),
        "squaring_alg_bytes": b_alg_sq,
        "squaring_hbm_frac": round(b_alg_sq / (ms_per_sq * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels_ms_per_squaring": {k: round(v, 5) for k, v in sorted(per_sq.items(), key=lambda kv: -kv[1])},
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT) if traffic is not None else None,
            "alg_bytes_per_launch": alg_bytes(dom, U),
            "avg_launch_ms": round(avg_ms, 5),
            "valu_util": valu_util,
        },
    }
    if cpu_res is not None and "ms_per_squaring" in cpu_res:
        line["cpu_baseline"] = {
            "value": round(cpu_res["ms_per_squaring"], 3),
            "unit": "ms/squaring",
            "cores": 1,
            "kind": cpu_res["kind"],
            "sample": f"{cpu_res['squarings']} squarings of the same N ({cpu_res['seconds']:.1f} s, after "
                      f"{cpu_res['init_s']:.0f} s init), {cpu_res['transform']}, one pinned host core",
        }
    elif cpu_res is not None:
        line["cpu_baseline"] = {"value": None, "error": cpu_res.get("error")}
    T.close()
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
