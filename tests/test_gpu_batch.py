"""Batched tests on the GPU: many Proth / Fermat tests of several length
classes through the engine's batch ABI with continuous admission, against the
oracle's protocol (GMP MMGW + mpz_powm) test by test; each result the same
alone and in the mix; the batch ABI's per-slot multipliers and its argument
checks."""
import ctypes
import random

import pytest

import annie_amd
import oracle
from tests.helpers import proth

pytestmark = pytest.mark.gpu


def _proth_mix(seed, count, lo, hi):
    rng = random.Random(seed)
    return [{"N": proth(rng.randrange(3, 1 << 20) | 1, rng.randrange(lo, hi)), "is_euler": True}
            for _ in range(count)]


def _check_against_oracle(tests, res):
    for t, r in zip(tests, res):
        print(f"case N bits {t['N'].bit_length()}: {r['message'][-60:]}", flush=True)
        o = oracle.check(t["N"], is_euler=t.get("is_euler", False), a_ref=t.get("a", 0))
        if r["status"] == "factor":
            continue  # gcd(N, R) > 1 for the engine's R: reported, not tested
        assert r["status"] == "success" and r["gl_ok"], r
        assert (r["res64"], r["is_prp"], r["a"]) == (o["res64"], o["is_prp"], o["a"]), (r, o)


def test_batch_proth_mix_matches_oracle(gpu):
    tests = _proth_mix(41, 24, 1500, 9000)
    tests += [{"N": proth(3, 2208), "is_euler": True}, {"N": proth(3, 2816), "is_euler": True}]
    res, st = annie_amd.batch_run(tests, engine="gpu", max_slots=5)
    print("stats", st, flush=True)
    assert st["rc"] == 0 and st["classes"] >= 2
    _check_against_oracle(tests, res)
    assert res[-2]["is_prp"] and res[-1]["is_prp"]


def test_batch_fermat_mix_matches_powm(gpu):
    rng = random.Random(42)
    tests = [{"N": rng.getrandbits(rng.randrange(8000, 40000)) | 1, "a": 3} for _ in range(8)]
    res, st = annie_amd.batch_run(tests, engine="gpu")
    for t, r in zip(tests, res):
        if r["status"] == "factor":
            continue
        assert r["status"] == "success" and r["gl_ok"]
        assert r["res64"] == oracle.powm(3, t["N"] - 1, t["N"]) & (2 ** 64 - 1)


@pytest.mark.parametrize("host_threads", [1, 4])
def test_graph_replays_match_launches(gpu, host_threads):
    tests = _proth_mix(45, 10, 1500, 5000)
    a, _ = annie_amd.batch_run(tests, engine="gpu", max_slots=3, host_threads=host_threads)
    b, _ = annie_amd.batch_run(tests, engine="gpu", max_slots=3, graphs=True, host_threads=host_threads)
    for x, y in zip(a, b):
        assert (x["status"], x["res64"], x["is_prp"], x["gl_ok"]) == (y["status"], y["res64"], y["is_prp"], y["gl_ok"]), \
            (x["message"], y["message"])


def test_same_result_alone_and_in_the_mix(gpu):
    mix = _proth_mix(43, 12, 2000, 6000)
    alone = [annie_amd.batch_run([t], engine="gpu")[0][0] for t in mix[:3]]
    plain = [annie_amd.check(t["N"], is_euler=True, gmp_threshold_digits=1, min_ln=5) for t in mix[:3]]
    for i, t in enumerate(mix):
        t["arrival"] = 0.01 * i
    res, _ = annie_amd.batch_run(mix, engine="gpu", max_slots=4, realtime=True)
    for a, p, r in zip(alone, plain, res[:3]):
        if r["status"] == "factor":
            assert a["status"] == "factor" and p["status"] == "factor"
            continue
        assert (a["res64"], a["is_prp"], a["gl_ok"]) == (r["res64"], r["is_prp"], r["gl_ok"])
        assert (p["res64"], p["is_prp"], p["gl_ok"]) == (r["res64"], r["is_prp"], r["gl_ok"])
    for r in res:
        assert r["admitted"] >= r["arrival"] - 1e-6


def _batch_for(Ns, a):
    lib = annie_amd.load_library()
    plans = {annie_amd.select_plan(N, a) for N in Ns}
    assert len(plans) == 1
    K, pck, LN, lbase = plans.pop()
    h = ctypes.c_void_p()
    assert lib.annie_hip_batch_create(K, pck, LN, lbase, len(Ns), 0, ctypes.byref(h)) == 0, annie_amd._err(lib)
    for s, N in enumerate(Ns):
        arr, n = annie_amd.int_to_limbs(N)
        f = (ctypes.c_uint64 * (n + 1))()
        nf = ctypes.c_size_t()
        rc = lib.annie_hip_batch_load(h, s, arr, n, a, f, n + 1, ctypes.byref(nf))
        assert rc == 0, annie_amd._err(lib)
    return lib, h


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_abi_per_slot_multipliers(gpu, graphs):
    """r0 = a_s * r0^2 per slot, with a different multiplier pattern in each
    slot: after t steps slot s holds 3^(its bit string) mod N_s.  With graphs
    the squarings are replays of one captured graph; the values are read back
    (eager work and a synchronisation) between groups of replays and checked
    at every read."""
    rng = random.Random(44)
    Ns = []
    while len(Ns) < 6:
        N = rng.getrandbits(30000) | (1 << 29999) | 1
        if annie_amd.select_plan(N, 3) == annie_amd.select_plan(Ns[0], 3) if Ns else True:
            t = annie_amd.Transform(N, 3, min_ln=5)
            ok = not t.init_failed()
            t.close()
            if ok:
                Ns.append(N)
    lib, h = _batch_for(Ns, 3)
    u32 = ctypes.c_uint32
    stride = (30000 // 64) + 4
    out = (ctypes.c_uint64 * (6 * stride))()
    nout = (ctypes.c_size_t * 6)()

    def values(ids, n):
        assert lib.annie_hip_batch_get(h, 0, ids, n, out, stride, nout) == 0, annie_amd._err(lib)
        return [sum(int(v) << (64 * i) for i, v in enumerate(out[j * stride:j * stride + nout[j]])) for j in range(n)]

    try:
        if graphs:
            assert lib.annie_hip_batch_set_graphs(h, 1) == 0, annie_amd._err(lib)
        ids = (u32 * 6)(*range(6))
        assert lib.annie_hip_batch_set1(h, ids, 6) == 0
        bits = [[rng.randrange(2) for _ in range(40)] for _ in range(6)]
        for step in range(40):
            a = (u32 * 6)(*[3 if bits[s][step] else 1 for s in range(6)])
            assert lib.annie_hip_batch_square_mul(h, ids, a, 6) == 0, annie_amd._err(lib)
            if step % 10 == 9:
                for s, v in enumerate(values(ids, 6)):
                    e = int("".join(map(str, bits[s][:step + 1])), 2)
                    assert v == pow(3, e, Ns[s]), f"slot {s} after {step + 1} steps"
        # a sub-list: only slots 4 and 1 take one more squaring
        sub = (u32 * 2)(4, 1)
        assert lib.annie_hip_batch_square_mul(h, sub, (u32 * 2)(3, 3), 2) == 0
        for s, v in enumerate(values(ids, 6)):
            e = int("".join(map(str, bits[s])), 2)
            if s in (1, 4):
                e = 2 * e + 1
            assert v == pow(3, e, Ns[s]), f"slot {s}"
        # snapshot flag: slot 2 squares with bit 31 set, then r1 = r1 * snapshot
        r0 = values(ids, 6)
        assert lib.annie_hip_batch_copy(h, 1, 0, ids, 6) == 0
        flag = (u32 * 6)(*[(1 | 0x80000000) if s == 2 else 1 for s in range(6)])
        assert lib.annie_hip_batch_square_mul(h, ids, flag, 6) == 0
        assert lib.annie_hip_batch_square_mul(h, ids, (u32 * 6)(*[1] * 6), 6) == 0
        one = (u32 * 1)(2)
        assert lib.annie_hip_batch_mul(h, 1, 3, one, 1) == 0, annie_amd._err(lib)
        assert lib.annie_hip_batch_get(h, 1, one, 1, out, stride, nout) == 0
        got = sum(int(v) << (64 * i) for i, v in enumerate(out[:nout[0]]))
        assert got == r0[2] * pow(r0[2], 2, Ns[2]) % Ns[2]
        # argument checks: duplicate slot, slot out of range, empty list
        assert lib.annie_hip_batch_set1(h, (u32 * 2)(1, 1), 2) != 0
        assert lib.annie_hip_batch_set1(h, (u32 * 1)(6), 1) != 0
        assert lib.annie_hip_batch_set1(h, ids, 0) != 0
    finally:
        lib.annie_hip_destroy(h)
