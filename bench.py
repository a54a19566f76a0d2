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
`--warmup` untimed ones (the negacyclic half of each squaring runs on a
second stream), after a GPU clock warm-up of `--gpu-warmup-seconds`.  The
residue is checked against a GMP fixture (tests/golden/cfg3_prefix.json,
mpz_powm): runs of at most 1,000 squarings in all (lead-in, warm-up, timed and
profile windows) end exactly at the fixture's 1,000-iteration prefix and
compare the final residue; longer runs compare the residue after its
100-iteration prefix, before the warm-up.  A mismatch prints an error line
and exits 1.  `roofline` is computed live for the kernel that takes the most
time per squaring (HIP event pairs around its launches over a profile window
of the same workload, issued on one stream so each kernel is timed alone),
against HBM and, with the VALU instruction count of the committed counter
summary, against the VALU peak; `cpu_baseline` is the reference's own CPU
transform (src/transformCPU.h, built from the reference sources into
oracle/_ref) timed on a bounded sample on one host core after the GPU work.
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


def alg_bytes(kind, U, K=5):
    """Algorithmic HBM bytes of one launch, U = K * n * 4 (one register).

    Every NTT pass reads and writes each job's register once (2U per job);
    the fused row pass of a product also reads the transformed operand (U).
    The carry reads the convolution residues (U) and writes a register (U),
    except that in the squaring's two-job carry the negacyclic job writes T_Q
    as three raw digit limbs (3U/K), which the last one-job carry reads back;
    the one-job carries (the m' carry and the last one) are timed under one
    name, so theirs is the mean of 2U and 2U + 3U/K.
    """
    name, jobs = kind.rsplit("x", 1)
    jobs = int(jobs)
    digits = 3 * U // K
    if name.startswith("k_row"):
        mode = name.split(",")[1].rstrip(">")
        return jobs * (3 * U if mode in ("mulc", "mulv") else 2 * U)
    if name.startswith("k_carry"):
        return 3 * U + digits if jobs == 2 else 2 * U + digits // 2
    return jobs * 2 * U


PMC_SUMMARY = os.path.join(ROOT, "profiles", "square_mul_40M_digits_pmc.md")
FIXTURE = os.path.join(ROOT, "tests", "golden", "cfg3_prefix.json")
# VALU peak: 256 CUs x 4 SIMDs x 16 lanes per cycle at 2.4 GHz (32-bit integer
# multiply-high included: it issues at full rate on gfx950).
VALU_PEAK = 256 * 64 * 2.4e9


def measured_counters(kernel):
    """Counters of one launch of `kernel` from the committed rocprofv3 summary
    of this workload: HBM bytes (FETCH_SIZE / WRITE_SIZE scaled by a plain copy
    of known size), VALU utilisation and VALU wave instructions per launch, or
    {} when the summary has no line for it."""
    try:
        lines = open(PMC_SUMMARY).read().splitlines()
    except OSError:
        return {}
    for ln in lines:
        f = ln.split()
        if len(f) < 2 or f[0] != "traffic" or f[1] != kernel:
            continue
        kv = dict(x.split("=", 1) for x in f[2:])
        b = (float(kv["fetch_kib"]) * float(kv["fetch_factor"]) + float(kv["write_kib"]) * float(kv["write_factor"])) * 1024
        out = {"traffic": int(round(b))}
        if "valu_util" in kv:
            out["valu_util"] = float(kv["valu_util"])
        if "valu_insts" in kv:
            out["valu_insts"] = float(kv["valu_insts"])
        return out
    return {}


def load_fixture(candidate):
    """{iterations: (res64, sha256)} of the GMP prefix fixture for the seeded
    40M-digit N (tests/golden/cfg3_prefix.json), or {} for another N."""
    try:
        fx = json.load(open(FIXTURE))
    except OSError:
        return {}
    if (fx["bits"], fx["seed"], fx["base"], fx["candidate"]) != (BITS, SEED, BASE, candidate):
        return {}
    return {e["iterations"]: (e["res64"], e["sha256"]) for e in fx["entries"]}


def verify(T, N, iterations, fixture):
    """Canonical residue of r0 after `iterations` loop iterations against the
    GMP fixture (SHA-256 of the whole residue and its low 64 bits)."""
    import hashlib
    v = T.get(0)
    res64, sha = fixture[iterations]
    digest = hashlib.sha256(v.to_bytes((N.bit_length() + 7) // 8, "little")).hexdigest()
    ok = (v & ((1 << 64) - 1)) == res64 and digest == sha
    progress(f"residue after {iterations} iterations {'matches' if ok else 'DIFFERS FROM'} the GMP fixture")
    return {"iterations": iterations, "res64": f"{v & ((1 << 64) - 1):016X}", "match": ok,
            "against": "mpz_powm, " + os.path.relpath(FIXTURE, ROOT)}


def host_info():
    """Host cores and CPU model of the machine the CPU baseline ran on."""
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"host_cores": os.cpu_count(), "host_cores_usable": usable, "cpu_model": model}


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


def jacobi(a, n):
    """Jacobi symbol (a/n), n odd positive."""
    a %= n
    r = 1
    while a:
        while a % 2 == 0:
            a //= 2
            if n % 8 in (3, 5):
                r = -r
        a, n = n, a
        if a % 4 == 3 and n % 4 == 3:
            r = -r
        a %= n
    return r if n == 1 else 0


def euler_base(N):
    """Base of the Euler (Proth) test: smallest a >= 2 with (a/N) = -1,
    skipping perfect squares (src/annie.h:249-253)."""
    a = 2
    while jacobi(a, N) != -1:
        a += 1
        s = math.isqrt(a)
        if s * s == a:
            a += 1
    return a


def gl_block(size):
    """Gerbicz-Li block B of an exponent of `size` bits (src/annie.h:95-101)."""
    lg = size.bit_length() - 1
    return (size - 1) // (2 << (lg // 2)) + 1


def protocol_squarings(N, is_euler):
    """square_mul operations of one whole test: the loop over e, the Euler
    residue step and the Gerbicz-Li verification (src/annie.h:319-404)."""
    e = (N - 1) >> 1 if is_euler else N - 1
    L = e.bit_length()
    B = gl_block(L)
    res = 0
    while e:
        res += e & ((1 << B) - 1)
        e >>= B
    return L + res.bit_length() + (1 if is_euler else 0)


def proth_batch_workload(count=4096, seed=4):
    """4,096 Proth candidates k*2^n+1: k odd uniform in [3, 2^20), n
    log-uniform in [2*10^4, 4*10^5] (seeded)."""
    rng = random.Random(seed)
    tests = []
    for _ in range(count):
        k = rng.randrange(3, 1 << 20) | 1
        n = int(math.exp(math.log(2e4) + rng.random() * math.log(20)))
        tests.append({"N": k * (1 << n) + 1, "is_euler": True, "k": k, "n": n})
    return tests


def plain_rate(t0, steps):
    """square_mul/s of the plain one-test engine on test t0 (HIP events over
    `steps` squarings of its exponent), or None when its transform does not
    initialise (gcd(N, R) > 1)."""
    import annie_amd
    T = annie_amd.Transform(t0["N"], t0["a"], min_ln=5)
    if T.init_failed():
        T.close()
        return None
    e = (t0["N"] - 1) >> 1
    nb = min(steps, e.bit_length())
    bits = [int(c) for c in format(e >> (e.bit_length() - nb), "b")]
    T.set1()
    T.square_mul_bits(bits[:20], t0["a"])
    ms = T.time_square_mul_bits(bits[20:], t0["a"])
    T.close()
    return (len(bits) - 20) / (ms * 1e-3)


def cpu_class_rate(job):
    """Child process: square_mul/s of the reference's own transform for one
    test on one core -- transformGMP below 10,000 decimal digits and
    transformCPU above, as annie::check chooses (src/annie.h:263-268)."""
    N, a, seconds = job
    import oracle  # the CPU baseline leg only
    kind = 0 if N.bit_length() < 33219 else 1
    ref = oracle.ref_available()
    T = oracle.RefTransform(N, a, kind) if ref else (oracle.GmpTransform(N, a) if kind == 0 else oracle.NttTransform(N, a))
    if T.init_failed():
        return None
    T.set1()
    done, t0 = 0, time.time()
    while done < 3 or time.time() - t0 < seconds:
        T.square_mul(a)
        done += 1
    dt = time.time() - t0
    T.close()
    return {"rate": done / dt, "kind": "reference" if ref else "port"}


def batch_roofline(cls_tests, steps, repeats=3):
    """Live roofline of the batched path on one class: a batch of the class's
    tests timed with HIP events around each launch (the engine's profile
    window) -- with steps > 0, `repeats` squaring-run launches of `steps`
    squarings per slot; with steps = 0, `repeats` x 8 batched squarings
    issued launch by launch, the dominant kernel reported.  Algorithmic bytes
    of a single-pass squaring: 11 U (U = K n 4) -- per launch of the launched
    path, alg_bytes() of the kernel times the slots -- the compulsory traffic
    of the registers (L2/MALL-resident at these lengths, so a bound, not HBM
    traffic)."""
    import ctypes

    import annie_amd
    lib = annie_amd.load_library()
    K, pck, LN, lbase = annie_amd.select_plan(cls_tests[0]["N"], cls_tests[0]["a"])
    h = ctypes.c_void_p()
    if lib.annie_hip_batch_create(K, pck, LN, lbase, len(cls_tests), 0, ctypes.byref(h)) != 0:
        return None
    ids = []
    for t in cls_tests:
        arr, n = annie_amd.int_to_limbs(t["N"])
        f = (ctypes.c_uint64 * (n + 1))()
        nf = ctypes.c_size_t()
        if lib.annie_hip_batch_load(h, len(ids), arr, n, t["a"], f, n + 1, ctypes.byref(nf)) == 0:
            ids.append(len(ids))
    try:
        u32 = ctypes.c_uint32
        sl = (u32 * len(ids))(*ids)
        lib.annie_hip_batch_set1(h, sl, len(ids))
        if steps > 0:
            mult = (u32 * (len(ids) * steps))(*([3] * (len(ids) * steps)))
            launch = lambda: lib.annie_hip_batch_square_run(h, sl, len(ids), mult, steps)  # noqa: E731
            per, sq = repeats, len(ids) * steps
        else:
            mult = (u32 * len(ids))(*([3] * len(ids)))
            launch = lambda: lib.annie_hip_batch_square_mul(h, sl, mult, len(ids))  # noqa: E731
            per, sq = repeats * 8, len(ids)
        launch()  # warm-up
        lib.annie_hip_batch_sync(h)
        lib.annie_hip_profile_begin(h)
        for _ in range(per):
            launch()
        buf = ctypes.create_string_buffer(1 << 14)
        lib.annie_hip_profile_end(h, buf, 1 << 14)
    finally:
        lib.annie_hip_destroy(h)
    prof = {}
    for ln in buf.value.decode().splitlines():
        name, cnt, tot = ln.split()
        prof[name] = (int(cnt), float(tot))
    name = max(prof, key=lambda k: prof[k][1])
    cnt, tot = prof[name]
    avg_ms = tot / cnt
    U = K * (1 << LN) * 4
    alg = len(ids) * (steps * 11 * U if steps > 0 else alg_bytes(name, U, K))
    achieved = alg / (avg_ms * 1e-3) / 1e9
    total_ms = sum(v[1] for v in prof.values())
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "alg_bytes_per_launch": alg,
            "avg_launch_ms": round(avg_ms, 4), "slots": len(ids), "plan": [K, pck, LN, lbase],
            "squarings_per_s_in_window": round(per * sq / (total_ms * 1e-3), 1)}


def bench_proth_batch(args):
    """Batched throughput on the 4,096-test Proth workload, measured: every
    test runs to completion through the batch driver (one engine batch per
    length class, continuous admission, squaring runs of up to --run-steps
    per test per launch), tests/hour = tests / wall time from the first
    admission to the last result.  Parity: the results of the tests in the
    committed GMP sample (tests/golden/config_res64.json, by workload index)
    must match.  Plain path and CPU baseline: per class, the measured
    square_mul rate of one test (the plain GPU engine; the reference's CPU
    transform on one core, in a pool of --cpu-procs processes), tests/hour =
    tests / sum over classes of (the class's square_mul count / rate) --
    `cores` processes for the CPU."""
    import multiprocessing as mp

    import annie_amd
    tests = proth_batch_workload(args.tests)
    classes = {}
    for t in tests:
        t["a"] = euler_base(t["N"])
        classes.setdefault(annie_amd.select_plan(t["N"], t["a"]), []).append(t)
    progress(f"{len(tests)} Proth tests in {len(classes)} length classes; batched run (run steps {args.run_steps})")
    batch = [{"N": t["N"], "is_euler": True} for t in tests]
    res, st = annie_amd.batch_run(batch, engine="gpu", max_slots=args.max_slots, run_steps=args.run_steps,
                                  run_min_active=args.run_min_active, host_threads=args.host_threads, verbose=True)
    wall = st["seconds"]
    done = sum(r["status"] in ("success", "factor") for r in res)
    bad = [r for r in res if r["status"] not in ("success", "factor")]
    tph_b = done / wall * 3600
    progress(f"batched: {done} of {len(tests)} tests in {wall:.1f} s ({st['setup_seconds']:.1f} s set-up), "
             f"{tph_b:.0f} tests/hour, {len(bad)} errors")
    verified = {"checked": 0, "match": 0, "against": "GMP mpz_powm, tests/golden/config_res64.json"}
    fx_path = os.path.join(ROOT, "tests", "golden", "config_res64.json")
    if args.tests == 4096 and os.path.exists(fx_path):
        for e in json.load(open(fx_path))["cfg4"]:
            r = res[e["index"]]
            ok = (r["status"] == "factor") if "factor" in e else (
                r["status"] == "success" and r["gl_ok"] and (r["res64"], r["is_prp"]) == (e["res64"], e["is_prp"]))
            verified["checked"] += 1
            verified["match"] += int(ok)
    if bad or verified["match"] != verified["checked"]:
        print(json.dumps({"error": "batched results differ from the GMP sample or failed", "verified": verified,
                          "errors": [r["message"] for r in bad[:5]]}), flush=True)
        sys.exit(1)

    # Plain GPU path and CPU baseline, per class.
    rows, hours_p, hours_c, jobs, order = [], 0.0, 0.0, [], []
    for key in sorted(classes, key=lambda k: (k[2], k[0])):
        ct = classes[key]
        t0 = next((t for t in ct if annie_amd.select_plan(t["N"], t["a"]) == key), None)
        rate_p = None
        for t in ct:
            rate_p = plain_rate(t, args.plain_steps)
            if rate_p:
                t0 = t
                break
        # tests answered at admission (gcd(N, R) > 1, R = B^n - 1) run no squarings
        R = (1 << ((key[1] * key[3]) << key[2])) - 1
        work = sum(protocol_squarings(t["N"], True) for t in ct if math.gcd(t["N"], R) == 1)
        rows.append({"K": key[0], "pck": key[1], "LN": key[2], "lbase": key[3], "tests": len(ct), "square_muls": work,
                     "plain_sq_per_s": round(rate_p, 1) if rate_p else None})
        if rate_p:
            hours_p += work / rate_p / 3600
        jobs.append((t0["N"], t0["a"], args.cpu_class_seconds))
        order.append(len(rows) - 1)
    tph_p = len(tests) / hours_p if hours_p else None
    cpu = None
    if not args.no_cpu_baseline:
        procs = max(1, min(args.cpu_procs, len(os.sched_getaffinity(0))))
        progress(f"CPU baseline: the reference's transform per class, {procs} processes")
        with mp.get_context("spawn").Pool(procs) as pool:
            rates = pool.map(cpu_class_rate, jobs, chunksize=1)
        for i, r in zip(order, rates):
            rows[i]["cpu_sq_per_s"] = round(r["rate"], 1) if r else None
            if r:
                hours_c += rows[i]["square_muls"] / r["rate"] / 3600
        kind = next((r["kind"] for r in rates if r), "reference")
        cpu = {"value": round(procs * len(tests) / hours_c, 1), "unit": "tests/hour", "cores": procs, "kind": kind,
               "sample": f"per class, one test's square_mul rate over {args.cpu_class_seconds} s on one core "
                         f"({procs} single-threaded processes at once), times the class's square_mul count; "
                         f"{procs} cores = {procs} independent processes", **host_info()}
    # Live roofline of the run kernel on the class with the most square_mul work.
    big = max(range(len(rows)), key=lambda i: rows[i]["square_muls"])
    big = (rows[big]["K"], rows[big]["pck"], rows[big]["LN"], rows[big]["lbase"])
    roof = batch_roofline(classes[big][:256], args.run_steps)
    return {
        "metric": "batched Proth tests/hour (4,096 k*2^n+1, n log-uniform in [2e4, 4e5]), measured completions",
        "value": round(tph_b, 1), "unit": "tests/hour", "n_gpus": 1, "higher_is_better": True,
        "seconds": round(wall, 2), "setup_seconds": round(st["setup_seconds"], 2), "completed": done,
        "plain_tests_per_hour": round(tph_p, 1) if tph_p else None,
        "speedup_vs_plain": round(tph_b / tph_p, 2) if tph_p else None,
        "plain_note": "plain = one test at a time on the plain GPU engine, per-class measured rate x work",
        "dtype": "u32", "data": "synthetic", "verified": verified,
        "config": {"workload": "4,096 Proth k*2^n+1, k odd in [3, 2^20), n log-uniform [2e4, 4e5], seed 4, Euler test",
                   "tests": len(tests), "classes": len(classes), "batches": st["classes"],
                   "run_steps": args.run_steps, "run_min_active": args.run_min_active,
                   "host_threads": args.host_threads, "max_slots": args.max_slots},
        "roofline": roof,
        "cpu_baseline": cpu,
        "classes": rows,
    }


def serve_request(rng, lo=1e4, hi=4e4):
    """One request of the serving mix: Fermat (random odd b-bit N, base 3) or
    Proth (k odd < 2^32, n = b - bits(k), Euler test) with probability 1/2,
    b log-uniform in [lo, hi] bits."""
    b = int(math.exp(math.log(lo) + rng.random() * math.log(hi / lo)))
    if rng.random() < 0.5:
        return {"N": rng.getrandbits(b) | (1 << (b - 1)) | 1, "a": 3}
    k = rng.randrange(3, 1 << 32) | 1
    return {"N": k * (1 << (b - k.bit_length())) + 1, "is_euler": True}


def serve_workload(rate, seconds, lo, hi, seed=5):
    """Poisson arrivals at `rate` requests/s over `seconds` (seeded), bits
    log-uniform in [lo, hi]."""
    rng = random.Random(seed)
    t, tests = 0.0, []
    while True:
        t += rng.expovariate(rate)
        if t > seconds:
            return tests
        r = serve_request(rng, lo, hi)
        r["arrival"] = t
        tests.append(r)


def gmp_res64(job):
    """Child process: (is_prp, res64) of a whole test from mpz_powm (Fermat
    a^(N-1); Euler a^((N-1)/2), then its square, src/annie.h:343-366)."""
    N, a, euler = job
    import oracle  # the checker only
    if euler:
        x = oracle.powm(a, (N - 1) >> 1, N)
        return x == N - 1, (x * x % N) & ((1 << 64) - 1)
    x = oracle.powm(a, N - 1, N)
    return x == 1, x & ((1 << 64) - 1)


def bench_serve(args):
    """Serving at the mixed-size configuration: requests (Fermat base 3 or
    Proth/Euler, 1/2 each, bits log-uniform in [1e4, --serve-max-bits]) arrive
    as a Poisson process for --serve-seconds at 70% of the capacity measured
    on a sample of the mix; each is admitted into a free slot of its length
    class as it arrives (continuous batching, squaring runs for classes with
    enough running tests) and runs to completion.  The run stops
    --serve-drain-seconds after the last arrival: requests still running then
    are reported as unfinished (at up to 2e7 bits a single test needs hours),
    latency quantiles are over completed requests.  Completed results are
    checked against mpz_powm on a sample (the smallest ones first)."""
    import multiprocessing as mp

    import annie_amd
    lo, hi = 1e4, args.serve_max_bits
    rng = random.Random(50)
    sample = [serve_request(rng, lo, hi) for _ in range(200)]
    kw = dict(max_slots=args.max_slots, run_steps=args.run_steps, run_min_active=args.run_min_active,
              host_threads=args.host_threads)
    cres, st = annie_amd.batch_run(sample, engine="gpu", time_limit=args.calib_seconds, **kw)
    errs = [r["message"] for r in cres if r["status"] == "error"]
    if st["rc"] != 0 or errs or st["squarings"] == 0:
        print(json.dumps({"error": "serving calibration failed", "stats": st, "errors": errs[:5]}), flush=True)
        sys.exit(1)
    cap_sq = st["squarings"] / max(st["seconds"] - st["setup_seconds"], 1e-9)
    mean_sq = sum(protocol_squarings(t["N"], t.get("is_euler", False)) for t in sample) / len(sample)
    rate = 0.7 * cap_sq / mean_sq
    tests = serve_workload(rate, args.serve_seconds, lo, hi)
    progress(f"serving: capacity {cap_sq:.0f} square_mul/s, mean {mean_sq:.0f} per request -> "
             f"{rate:.3f} tests/s offered, {len(tests)} requests over {args.serve_seconds} s")
    res, st2 = annie_amd.batch_run(tests, engine="gpu", realtime=True, verbose=True,
                                   time_limit=args.serve_seconds + args.serve_drain_seconds, **kw)
    ok = [r for r in res if r["status"] in ("success", "factor")]
    unfinished = [r for r in res if r["status"] == "aborted"]
    bad = [r for r in res if r["status"] not in ("success", "factor", "aborted")]
    tested = [r for r in ok if r["status"] == "success"]
    lat = sorted(r["finished"] - r["arrival"] for r in ok)
    lat_t = sorted(r["finished"] - r["arrival"] for r in tested)
    # GMP check of completed results, smallest first
    done = sorted((i for i, r in enumerate(res) if r["status"] == "success"), key=lambda i: tests[i]["N"].bit_length())
    chk = done[:args.serve_check]
    jobs = [(tests[i]["N"], res[i]["a"], bool(tests[i].get("is_euler"))) for i in chk]
    with mp.get_context("spawn").Pool(max(1, min(16, len(os.sched_getaffinity(0))))) as pool:
        want = pool.map(gmp_res64, jobs, chunksize=1)
    match = sum((res[i]["is_prp"], res[i]["res64"]) == w for i, w in zip(chk, want))
    verified = {"checked": len(chk), "match": match, "largest_bits": max((tests[i]["N"].bit_length() for i in chk),
                                                                        default=0), "against": "GMP mpz_powm"}
    if bad or match != len(chk):
        print(json.dumps({"error": "served results failed or differ from GMP", "verified": verified,
                          "errors": [r["message"] for r in bad[:5]]}), flush=True)
        sys.exit(1)

    def q(v, p):
        return round(v[min(len(v) - 1, int(p * len(v)))], 3) if v else None
    window = args.serve_seconds + args.serve_drain_seconds
    return {
        "metric": f"served tests/hour with p50/p99 time-to-result (mixed Fermat/Proth, 1e4-{hi:.0e} bits)",
        "latency_note": "p50/p99/max over completed requests that ran the test; *_all_s also count requests "
                        "answered at admission with a factor of R (gcd(N, R) > 1); unfinished requests were "
                        "still running when the window closed",
        "value": round(len(ok) / window * 3600, 1), "unit": "tests/hour", "n_gpus": 1, "higher_is_better": True,
        "p50_s": q(lat_t, 0.50), "p99_s": q(lat_t, 0.99), "max_s": q(lat_t, 1.0),
        "p50_all_s": q(lat, 0.50), "p99_all_s": q(lat, 0.99),
        "offered_tests_per_s": round(rate, 4), "requests": len(tests), "completed": len(ok), "tested": len(tested),
        "factor_found": len(ok) - len(tested), "unfinished": len(unfinished),
        "unfinished_max_bits": max((tests[i]["N"].bit_length() for i, r in enumerate(res) if r["status"] == "aborted"),
                                   default=0),
        "largest_completed_bits": max((tests[i]["N"].bit_length() for i in done), default=0),
        "errors": len(bad), "verified": verified,
        "capacity_square_mul_per_s": round(cap_sq, 1), "classes": st2["classes"], "slots": st2["slots"],
        "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"Poisson arrivals, 70% of measured capacity, Fermat base 3 or Proth (Euler) 1/2 each, "
                               f"b log-uniform in [1e4, {hi:.0e}] bits, seed 5", "serve_seconds": args.serve_seconds,
                   "drain_seconds": args.serve_drain_seconds, "run_steps": args.run_steps,
                   "run_min_active": args.run_min_active},
    }


def bench_fermat332k(args):
    """The whole Fermat PRP test (base 3, Gerbicz-Li included) of the seeded
    ~332,000-bit N on the GPU path (plan K = 4, 2^13), seconds per test, its
    residue against the GMP fixture; the dominant kernel's live roofline over
    a profile window of the same squaring; the reference's own CPU transform
    (transformCPU, src/transformCPU.h, K = 5) timed on a bounded sample of the
    same N's squarings on one core, extrapolated to the test's square_mul
    count."""
    import annie_amd
    import oracle  # the CPU baseline leg and the fixture's N only
    from tests.golden.make_config_fixtures import cfg1_N
    N, cand = cfg1_N()
    fx_path = os.path.join(ROOT, "tests", "golden", "config_res64.json")
    fx = json.load(open(fx_path))["cfg1"] if os.path.exists(fx_path) else None
    progress("GPU: whole Fermat test of the 332,000-bit N")
    r = annie_amd.check(N, a=3, gmp_threshold_digits=1, min_ln=5)
    if r["status"] != "success" or not r["gl_ok"]:
        print(json.dumps({"error": r["message"]}), flush=True)
        sys.exit(1)
    verified = []
    if fx and fx.get("candidate") == cand:
        ok = (r["res64"], r["is_prp"]) == (fx["res64"], fx["is_prp"])
        verified.append({"res64": f"{r['res64']:016X}", "match": ok, "against": "mpz_powm, " +
                         os.path.relpath(fx_path, ROOT)})
        if not ok:
            print(json.dumps({"error": "GPU residue differs from the GMP fixture", "verified": verified}), flush=True)
            sys.exit(1)
    work = protocol_squarings(N, False)
    # profile window on the plain engine: per-kernel HIP event pairs
    T = annie_amd.Transform(N, 3, min_ln=5)
    U = T.K * T.n * 4
    bits = exponent_bits(N, 400)
    T.set1()
    T.square_mul_bits(bits[:100], 3)
    T.profile_begin()
    T.square_mul_bits(bits[100:400], 3)
    prof = T.profile_end()
    plan = (T.K, T.pck, T.LN, T.lbase)
    T.close()
    dom = max(prof, key=lambda k: prof[k][1])
    cnt, tot = prof[dom]
    achieved = alg_bytes(dom, U, plan[0]) / (tot / cnt * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        kind = "reference" if oracle.ref_available() else "port"
        C = oracle.RefTransform(N, 3, 1) if kind == "reference" else oracle.NttTransform(N, 3)
        C.set1()
        done, t0 = 0, time.time()
        while done < 3 or time.time() - t0 < args.cpu_seconds:
            C.square_mul(3 if bits[done % len(bits)] else 1)
            done += 1
        dt = time.time() - t0
        C.close()
        cpu = {"value": round(dt / done * work, 1), "unit": "s/test", "cores": 1, "kind": kind,
               "sample": f"{done} squarings of the same N ({dt:.1f} s) on one core with the reference's "
                         f"transformCPU, times the test's {work} square_mul operations",
               "ms_per_squaring": round(dt * 1e3 / done, 4), **host_info()}
    return {
        "metric": "s per Fermat PRP test, 332,000-bit N (Gerbicz-Li included)", "value": round(r["seconds"], 3),
        "unit": "s/test", "n_gpus": 1, "higher_is_better": False, "dtype": "u32", "data": "synthetic",
        "ms_per_squaring": round(r["seconds"] * 1e3 / work, 5), "square_muls": work, "verified": verified,
        "config": {"workload": f"Fermat PRP test, seeded random odd 332,000-bit N (seed 1, candidate {cand}), base 3",
                   "K": plan[0], "pck": plan[1], "LN": plan[2], "lbase": plan[3]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "alg_bytes_per_launch": alg_bytes(dom, U, plan[0]), "avg_launch_ms": round(tot / cnt, 5),
                     "note": "a one-pass length: each launch is a few microseconds, the squaring is bound by "
                             "launch latency, not by HBM or VALU"},
        "kernels_ms_per_squaring": {k: round(v[1] / 300, 5) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])},
        "cpu_baseline": cpu,
    }


def bench_proth3m(args):
    """ms/squaring of the Euler (Proth) test of k*2^3,000,000+1 (seeded 32-bit
    odd k), the reference's plan K = 4, LN = 16; the reference CPU transform
    timed on a bounded sample of the same N on one core.  After the timed
    squarings the run continues (untimed) to the next iteration count of the
    GMP fixture tests/golden/cfg2_proth.json and compares the residue."""
    import annie_amd
    import oracle  # the CPU baseline leg only
    rng = random.Random(2)
    while True:
        k = rng.getrandbits(32) | 1 | (1 << 31)
        N = k * (1 << 3_000_000) + 1
        a = euler_base(N)
        T = annie_amd.Transform(N, a)
        if not T.init_failed():
            break
        T.close()
    e = (N - 1) >> 1
    total = args.warmup + args.steps
    fx_path = os.path.join(ROOT, "tests", "golden", "cfg2_proth.json")
    check_at = None
    if os.path.exists(fx_path):
        fx = json.load(open(fx_path))
        if fx["k"] == k:
            # iteration bits(k) + j holds a^(k 2^j)
            pts = sorted((k.bit_length() + p["j"], p["low64"]) for p in fx["prefix"])
            check_at = next((p for p in pts if p[0] >= total), None)
    nbits = max(total, check_at[0] if check_at else 0)
    bits = [int(c) for c in format(e >> (e.bit_length() - nbits), "b")]
    T.set1()
    T.square_mul_bits(bits[:args.warmup], a)
    T.sync()
    ms = T.time_square_mul_bits(bits[args.warmup:total], a) / args.steps
    plan = (T.K, T.pck, T.LN, T.lbase)
    verified = []
    if check_at:
        T.square_mul_bits(bits[total:check_at[0]], a)
        low = T.get(0) & ((1 << 64) - 1)
        verified.append({"iterations": check_at[0], "res64": f"{low:016X}", "match": low == check_at[1],
                         "against": "GMP, " + os.path.relpath(fx_path, ROOT)})
    T.close()
    if any(not v["match"] for v in verified):
        print(json.dumps({"error": "GPU residue differs from the GMP fixture", "verified": verified}), flush=True)
        sys.exit(1)
    progress(f"GPU {ms:.4f} ms/squaring; CPU baseline (reference transformCPU, one core)")
    cpu = None
    if not args.no_cpu_baseline:
        kind = "reference" if oracle.ref_available() else "port"
        C = oracle.RefTransform(N, a, 1) if kind == "reference" else oracle.NttTransform(N, a)
        C.set1()
        done, t0 = 0, time.time()
        while done < len(bits) and (done < 2 or time.time() - t0 < args.cpu_seconds / 2):
            C.square_mul(a if bits[done] else 1)
            done += 1
        dt = time.time() - t0
        cpu = {"value": round(dt * 1e3 / done, 3), "unit": "ms/squaring", "cores": 1, "kind": kind,
               "sample": f"{done} squarings of the same N ({dt:.1f} s), one core", **host_info()}
        C.close()
    return {
        "metric": "ms/squaring, Proth k*2^3,000,000+1 (Euler test)", "value": round(ms, 5),
        "unit": "ms/squaring", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": False, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"Euler test of k*2^3000000+1, k = {k} (seed 2), base {a}",
                   "K": plan[0], "pck": plan[1], "LN": plan[2], "lbase": plan[3]},
        "verified": verified,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=800)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--gpu-warmup-seconds", type=float, default=1.0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fixture-check", action="store_true",
                    help="squaring: skip the residue check against the GMP fixture and its untimed lead-in "
                         "(profiler runs, where every dispatch is slow)")
    ap.add_argument("--cpu-baseline-worker", type=int, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--workload", choices=["squaring", "fermat332k", "proth3m", "proth-batch", "serve"],
                    default="squaring",
                    help="squaring: the ~40M-digit square_mul loop (default); proth3m: k*2^3,000,000+1; "
                         "proth-batch: 4,096 Proth tests, batched vs plain; serve: Poisson arrivals with "
                         "continuous batching")
    ap.add_argument("--tests", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=400)
    ap.add_argument("--max-slots", type=int, default=0)
    ap.add_argument("--run-steps", type=int, default=0,
                    help="proth-batch/serve: squarings per test per launch (squaring runs; 0 = off)")
    ap.add_argument("--run-min-active", type=int, default=4,
                    help="proth-batch/serve: running tests a class needs for squaring runs")
    ap.add_argument("--host-threads", type=int, default=0,
                    help="proth-batch/serve: host workers driving the classes (0 = the batch driver's default)")
    ap.add_argument("--cpu-procs", type=int, default=16)
    ap.add_argument("--cpu-class-seconds", type=float, default=1.0)
    ap.add_argument("--serve-seconds", type=float, default=120.0)
    ap.add_argument("--serve-drain-seconds", type=float, default=60.0)
    ap.add_argument("--serve-max-bits", type=float, default=2e7)
    ap.add_argument("--serve-check", type=int, default=32, help="serve: completed results checked against GMP")
    ap.add_argument("--calib-seconds", type=float, default=5.0)
    args = ap.parse_args()

    if args.cpu_baseline_worker is not None:
        cpu_baseline_worker(args.cpu_baseline_worker, args.cpu_seconds)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the path runs one test per GPU and this build has no "
              f"multi-GPU mode (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    if args.workload != "squaring":
        line = {"fermat332k": bench_fermat332k, "proth3m": bench_proth3m, "proth-batch": bench_proth_batch,
                "serve": bench_serve}[args.workload](args)
        if rank == 0:
            print(json.dumps(line), flush=True)
        return

    import annie_amd

    # Pick the first seeded N the engine accepts (gcd(N, R) = 1).
    skip = 0
    while True:
        N = workload_N(skip=skip)
        progress(f"candidate {skip}: {N.bit_length()}-bit N; creating the context (N^-1 mod R takes about a minute)")
        t0 = time.time()
        T = annie_amd.Transform(N, BASE)
        init_s = time.time() - t0
        progress(f"context ready in {init_s:.1f} s (init_failed={T.init_failed()})")
        if not T.init_failed():
            break
        T.close()
        skip += 1

    U = T.K * T.n * 4
    fixture = load_fixture(skip)
    total = args.warmup + args.steps + args.profile_steps
    # Residue check against the GMP fixture: when the whole run fits in the
    # fixture's longest prefix, untimed lead-in squarings bring the run to
    # exactly that many iterations and the final residue is compared; longer
    # runs compare the residue after the fixture's shortest prefix, taken
    # before the warm-up.
    if args.no_fixture_check:  # profiler runs: no lead-in, no residue check
        fixture = {}
    longest = max(fixture) if fixture else 0
    if fixture and total <= longest:
        lead, pre = longest - total, 0
    else:
        lead, pre = 0, (min(fixture) if fixture else 0)
    bits = exponent_bits(N, pre + lead + total)
    verified = []
    T.set1()
    if pre:
        progress(f"residue check: {pre} squarings against the GMP fixture")
        T.square_mul_bits(bits[:pre], BASE)
        verified.append(verify(T, N, pre, fixture))
    pos = pre
    if lead:
        progress(f"lead-in: {lead} untimed squarings")
        T.square_mul_bits(bits[pos:pos + lead], BASE)
        pos += lead
    # Clock warm-up: products on register 2 (r0 and r1 untouched) until the
    # GPU has been busy for --gpu-warmup-seconds.
    T.copy(2, 0)
    w0 = time.time()
    while time.time() - w0 < args.gpu_warmup_seconds:
        for _ in range(20):
            T.mul(2, 2)
        T.sync()
    progress(f"warmup: {args.warmup} squarings (after {time.time() - w0:.1f} s of GPU clock warm-up)")
    T.square_mul_bits(bits[pos:pos + args.warmup], BASE)
    pos += args.warmup
    T.sync()
    progress(f"timed region: {args.steps} squarings")
    w0 = time.time()
    ms = T.time_square_mul_bits(bits[pos:pos + args.steps], BASE)
    wall = time.time() - w0
    pos += args.steps
    ms_per_sq = ms / args.steps

    progress(f"{ms_per_sq:.4f} ms/squaring; profile window: {args.profile_steps} squarings")
    T.profile_begin()
    T.square_mul_bits(bits[pos:pos + args.profile_steps], BASE)
    prof = T.profile_end()
    pos += args.profile_steps
    if fixture and pos in fixture:
        verified.append(verify(T, N, pos, fixture))
    per_sq = {k: tot / args.profile_steps for k, (cnt, tot) in prof.items()}
    dom = max(prof, key=lambda k: prof[k][1])
    cnt, tot = prof[dom]
    avg_ms = tot / cnt
    achieved = alg_bytes(dom, U, T.K) / (avg_ms * 1e-3) / 1e9
    b_alg_sq = 27 * U  # (16P - 5) U with P = 2 two-pass directions at this length
    pmc = measured_counters(dom)

    # A plain register copy of known size (U bytes read, U written) for
    # calibrating HBM byte counters in profiling runs.
    T.copy(1, 0)
    T.sync()
    plan = (T.K, T.n, T.LN, T.lbase, T.pck)
    T.close()

    if any(not v["match"] for v in verified):
        print(json.dumps({"error": "GPU residue differs from the GMP fixture", "verified": verified}), flush=True)
        sys.exit(1)

    cpu_res = None
    if rank == 0 and not args.no_cpu_baseline:
        # After the GPU work, so that it competes with nothing that is timed.
        progress("CPU baseline (reference transformCPU on one core)")
        cpu = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", str(skip),
                              "--cpu-seconds", str(args.cpu_seconds)], capture_output=True, text=True, timeout=900)
        try:
            cpu_res = json.loads(cpu.stdout.strip().splitlines()[-1])
        except (ValueError, IndexError):
            cpu_res = {"error": (cpu.stderr or cpu.stdout)[-300:]}

    roof = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": pmc.get("traffic"),
        "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT) if pmc else None,
        "alg_bytes_per_launch": alg_bytes(dom, U, T.K),
        "avg_launch_ms": round(avg_ms, 5),
    }
    if pmc.get("valu_insts"):
        # VALU roof of the same launch: its counted VALU wave instructions
        # (64 lanes each) over its live duration, against 256 CUs x 64 lanes
        # per cycle at 2.4 GHz.
        rate = pmc["valu_insts"] * 64 / (avg_ms * 1e-3)
        roof["valu"] = {"achieved": round(rate / 1e12, 2), "peak": VALU_PEAK / 1e12, "unit": "T lane-ops/s",
                        "frac": round(rate / VALU_PEAK, 4), "counter_util": pmc.get("valu_util")}
        roof["limiter"] = "valu" if rate / VALU_PEAK > achieved / HBM_PEAK_GBS else "hbm"
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
            "bits": BITS, "seed": SEED, "candidate": skip, "K": plan[0], "n": plan[1], "LN": plan[2],
            "lbase": plan[3], "pck": plan[4], "parallelism": "single GPU",
        },
        "verified": verified,
        "gpu_warmup_s": args.gpu_warmup_seconds,
        "squarings_per_s": round(1e3 / ms_per_sq, 1),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 5),
        "init_s": round(init_s, 1),
        "squaring_alg_bytes": b_alg_sq,
        "squaring_hbm_frac": round(b_alg_sq / (ms_per_sq * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels_ms_per_squaring": {k: round(v, 5) for k, v in sorted(per_sq.items(), key=lambda kv: -kv[1])},
        "roofline": roof,
    }
    host = host_info()
    if cpu_res is not None and "ms_per_squaring" in cpu_res:
        line["cpu_baseline"] = {
            "value": round(cpu_res["ms_per_squaring"], 3),
            "unit": "ms/squaring",
            "cores": 1,
            "kind": cpu_res["kind"],
            "sample": f"{cpu_res['squarings']} squarings of the same N ({cpu_res['seconds']:.1f} s, after "
                      f"{cpu_res['init_s']:.0f} s init), {cpu_res['transform']}, one pinned host core",
            **host,
        }
    elif cpu_res is not None:
        line["cpu_baseline"] = {"value": None, "error": cpu_res.get("error"), **host}
    if rank == 0:
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
