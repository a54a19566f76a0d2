"""Squaring runs (annie_hip_batch_square_run: one workgroup per slot runs
whole squarings, many per launch) against the same squarings issued one by
one through annie_hip_batch_square_mul, for every supported class shape
(transform lengths 2^9 .. 2^13, K = 3, 4, 5): equal canonical values after
uneven runs, snapshots on the last squaring of a run, exact powers; and the
batch driver in run mode against the oracle's protocol."""
import ctypes
import math
import random

import pytest

import annie_amd
import oracle
from tests.helpers import proth

pytestmark = pytest.mark.gpu
u32 = ctypes.c_uint32


def _moduli(K, LN, count, seed):
    """`count` odd N of plan (K, *, LN, *) for base 3 whose transform initialises."""
    rng = random.Random(seed)
    lo, hi = 1000, 2_000_000
    while True:  # a bit size inside the class: bisect on LN, then on K
        mid = (lo + hi) // 2
        p = oracle.plan((1 << (mid - 1)) | 1, 3)
        if (p[2], p[0]) == (LN, K):
            break
        if (p[2], p[0]) < (LN, K):
            lo = mid + 1
        else:
            hi = mid - 1
        assert lo <= hi, f"no size for K={K}, LN={LN}"
    out = []
    while len(out) < count:
        N = rng.getrandbits(mid) | (1 << (mid - 1)) | 1
        if oracle.plan(N, 3)[:3:2] != (K, LN):
            continue
        t = annie_amd.Transform(N, 3, min_ln=5)
        ok = not t.init_failed()
        t.close()
        if ok:
            out.append(N)
    return out


def _batch(Ns):
    lib = annie_amd.load_library()
    K, pck, LN, lbase = annie_amd.select_plan(Ns[0], 3)
    h = ctypes.c_void_p()
    assert lib.annie_hip_batch_create(K, pck, LN, lbase, len(Ns), 0, ctypes.byref(h)) == 0, annie_amd._err(lib)
    for s, N in enumerate(Ns):
        arr, n = annie_amd.int_to_limbs(N)
        f = (ctypes.c_uint64 * (n + 1))()
        nf = ctypes.c_size_t()
        assert lib.annie_hip_batch_load(h, s, arr, n, 3, f, n + 1, ctypes.byref(nf)) == 0, annie_amd._err(lib)
    return lib, h


def _values(lib, h, reg, ids, Ns):
    stride = max(N.bit_length() for N in Ns) // 64 + 3
    out = (ctypes.c_uint64 * (len(ids) * stride))()
    nout = (ctypes.c_size_t * len(ids))()
    arr = (u32 * len(ids))(*ids)
    assert lib.annie_hip_batch_get(h, reg, arr, len(ids), out, stride, nout) == 0, annie_amd._err(lib)
    return [sum(int(v) << (64 * i) for i, v in enumerate(out[j * stride:j * stride + nout[j]])) for j in range(len(ids))]


CLASSES = [(3, 9), (4, 9), (5, 9), (3, 10), (4, 10), (5, 10), (3, 11), (4, 11), (5, 11), (3, 12), (4, 12),
           (5, 12), (3, 13), (4, 13), (5, 13)]


@pytest.mark.parametrize("K,LN", CLASSES)
def test_runs_match_one_by_one(gpu, K, LN):
    Ns = _moduli(K, LN, 3, 100 * K + LN)
    print(f"case runs K={K} LN={LN} bits={[N.bit_length() for N in Ns]}", flush=True)
    lib_r, hr = _batch(Ns)
    lib, h1 = _batch(Ns)
    assert lib.annie_hip_batch_run_supported(hr) == 1
    rng = random.Random(K * 31 + LN)
    ids = list(range(3))
    try:
        for hh in (hr, h1):
            assert lib.annie_hip_batch_set1(hh, (u32 * 3)(*ids), 3) == 0
            assert lib.annie_hip_batch_copy(hh, 1, 0, (u32 * 3)(*ids), 3) == 0
        exps = [0, 0, 0]
        steps = 12
        for rnd in range(3):
            # uneven run lengths, slot order permuted, slot 2's last squaring snapshots in round 1
            order = rng.sample(ids, 3)
            lens = [rng.randrange(1, steps + 1) for _ in ids]
            mult = []
            for s in order:
                row = [3 if rng.getrandbits(1) else 1 for _ in range(lens[s])]
                if rnd == 1 and s == 2:
                    row[-1] |= 0x80000000
                for m in row:
                    exps[s] = 2 * exps[s] + (1 if (m & 0x7fffffff) == 3 else 0)
                    assert lib.annie_hip_batch_square_mul(h1, (u32 * 1)(s), (u32 * 1)(m), 1) == 0
                mult += row + [0] * (steps - len(row))
            rc = lib.annie_hip_batch_square_run(hr, (u32 * 3)(*order), 3, (u32 * len(mult))(*mult), steps)
            assert rc == 0, annie_amd._err(lib)
            got = _values(lib, hr, 0, ids, Ns)
            want = _values(lib, h1, 0, ids, Ns)
            assert got == want, f"round {rnd}"
            assert got == [pow(3, e, N) for e, N in zip(exps, Ns)]
            if rnd == 1:  # r1 *= snapshot (slot 2)
                for hh in (hr, h1):
                    assert lib.annie_hip_batch_mul(hh, 1, 3, (u32 * 1)(2), 1) == 0
                assert _values(lib, hr, 1, [2], Ns) == _values(lib, h1, 1, [2], Ns)
        # rejected: a snapshot before the end of a run, a multiplier after a zero
        bad = (u32 * 4)(3 | 0x80000000, 3, 0, 0)
        assert lib.annie_hip_batch_square_run(hr, (u32 * 1)(0), 1, bad, 4) == annie_amd.ANNIE_ERR_ARG
        bad = (u32 * 4)(3, 0, 3, 0)
        assert lib.annie_hip_batch_square_run(hr, (u32 * 1)(0), 1, bad, 4) == annie_amd.ANNIE_ERR_ARG
    finally:
        lib.annie_hip_destroy(hr)
        lib.annie_hip_destroy(h1)


def test_driver_run_mode_matches_oracle(gpu):
    """Proth tests of several classes through the batch driver in run mode."""
    rng = random.Random(61)
    tests = [{"N": proth(rng.randrange(3, 1 << 20) | 1, rng.randrange(12000, 60000)), "is_euler": True}
             for _ in range(24)]
    res, st = annie_amd.batch_run(tests, engine="gpu", run_steps=32, run_min_active=1)
    print("stats", st, flush=True)
    assert st["rc"] == 0
    for t, r in zip(tests, res):
        K, pck, LN, lbase = annie_amd.select_plan(t["N"], r["a"])
        g = math.gcd(t["N"], (1 << ((pck * lbase) << LN)) - 1)
        if g > 1:  # the engine's R = B^n - 1 shares a factor with N: reported, not tested
            assert r["status"] == "factor" and r["message"].startswith(str(g)), r
            continue
        o = oracle.check(t["N"], is_euler=True)
        assert r["status"] == "success" and r["gl_ok"], r
        assert (r["res64"], r["is_prp"], r["a"]) == (o["res64"], o["is_prp"], o["a"])


@pytest.mark.parametrize("K,LN,first,second", [(5, 11, (110750, 13, 17), (114750, 2, 18)),
                                               (4, 13, (311500, 5, 27), (327750, 2, 28))])
def test_batch_slots_of_two_lbase(gpu, K, LN, first, second):
    """Moduli of one (K, LN) whose plans differ in lbase (the base changes the
    CRT bound, src/transformGPU.h:825-850) share one batch, each slot with
    its own digit width: squarings issued one by one and as runs give the
    exact powers a_s^e mod N_s."""
    rng = random.Random(K + LN)
    Ns, bases = [], []
    for bits, a, lbase in (first, second, first, second):
        while True:
            N = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
            p = annie_amd.select_plan(N, a)
            if (p[0], p[2], p[3]) != (K, LN, lbase):
                continue
            t = annie_amd.Transform(N, a, min_ln=5)
            ok = not t.init_failed()
            t.close()
            if ok:
                break
        Ns.append(N)
        bases.append(a)
    print(f"case mixed lbase K={K} LN={LN} lbase {first[2]}/{second[2]}", flush=True)
    lib = annie_amd.load_library()
    K0, pck, LN0, lb0 = annie_amd.select_plan(Ns[0], bases[0])
    h = ctypes.c_void_p()
    assert lib.annie_hip_batch_create(K0, pck, LN0, lb0, 4, 0, ctypes.byref(h)) == 0, annie_amd._err(lib)
    try:
        for s, (N, a) in enumerate(zip(Ns, bases)):
            arr, n = annie_amd.int_to_limbs(N)
            f = (ctypes.c_uint64 * (n + 1))()
            nf = ctypes.c_size_t()
            assert lib.annie_hip_batch_load(h, s, arr, n, a, f, n + 1, ctypes.byref(nf)) == 0, annie_amd._err(lib)
        ids = list(range(4))
        assert lib.annie_hip_batch_set1(h, (u32 * 4)(*ids), 4) == 0
        exps = [0] * 4
        for step in range(10):  # one by one
            m = [bases[s] if rng.getrandbits(1) else 1 for s in ids]
            for s in ids:
                exps[s] = 2 * exps[s] + (1 if m[s] != 1 else 0)
            assert lib.annie_hip_batch_square_mul(h, (u32 * 4)(*ids), (u32 * 4)(*m), 4) == 0
        steps = 9
        mult = []
        for s in ids:  # one run each
            row = [bases[s] if rng.getrandbits(1) else 1 for _ in range(steps)]
            for x in row:
                exps[s] = 2 * exps[s] + (1 if x != 1 else 0)
            mult += row
        assert lib.annie_hip_batch_square_run(h, (u32 * 4)(*ids), 4, (u32 * len(mult))(*mult), steps) == 0, \
            annie_amd._err(lib)
        got = _values(lib, h, 0, ids, Ns)
        assert got == [pow(a, e, N) for a, e, N in zip(bases, exps, Ns)]
    finally:
        lib.annie_hip_destroy(h)
