"""NTT launch schedules on the GPU: every schedule that the autotuner may pick
(annie_amd.schedules) gives the same registers as the default one, word for
word, on the two-pass lengths where the passes have knobs; the batch ABI with
a schedule set (launched and graph-replayed) gives a^e mod N exactly; the
engine refuses schedules that do not fit."""
import ctypes
import random

import pytest

import annie_amd
from tests.helpers import assert_same_words, usable_N

pytestmark = pytest.mark.gpu

# (bits, seed) -> two-pass lengths 2^14, 2^16 (the Proth k*2^3,000,000+1
# class, K = 4), 2^18 and 2^19 (columns of 2^9 at both).
CASES = [(1000000, 5, 14), (3000020, 11, 16), (16000000, 13, 18), (30000000, 14, 19)]


@pytest.mark.parametrize("bits,seed,LN", CASES)
def test_every_schedule_gives_the_default_registers(gpu, bits, seed, LN):
    a = 3
    N = usable_N(bits, seed, a)
    t = annie_amd.Transform(N, a, min_ln=5)
    try:
        assert t.LN == LN
        scheds = annie_amd.schedules(LN)
        assert len(scheds) >= 4
        print(f"case bits={bits} LN={LN}: {len(scheds)} schedules", flush=True)
        bits_e = [1, 0, 1, 1, 0, 1, 1, 1]
        e = int("".join(map(str, bits_e)), 2)
        ref = None
        for sch in scheds:
            t.set_schedule(*sch)
            assert t.get_schedule() == (sch if sch != (0, 0) else annie_amd.schedule_default(LN))
            t.set1()
            for b in bits_e:
                t.square_mul(a if b else 1)
            words = t.read_register(0)
            if ref is None:
                ref = words
            else:
                assert_same_words(words, ref, f"r0 with schedule {sch}", t.n)
            assert t.get(0) == pow(a, e, N), sch
        t.set_schedule(0, 0)
    finally:
        t.close()


def test_schedules_that_do_not_fit_are_refused(gpu):
    N = usable_N(1000000, 5, 3)
    t = annie_amd.Transform(N, 3, min_ln=5)
    try:
        assert t.LN == 14
        for bad in [(3, 0), (0, 5), (128, 0), (0, 128), (-1, 0)]:
            with pytest.raises(RuntimeError):
                t.set_schedule(*bad)
        assert t.get_schedule() == annie_amd.schedule_default(14)
    finally:
        t.close()
    N = usable_N(300000, 6, 3)
    t = annie_amd.Transform(N, 3, min_ln=5)
    try:
        assert t.LN == 13
        with pytest.raises(RuntimeError):
            t.set_schedule(0, 8)  # single-pass length: no column pass
        assert t.get_schedule() == (0, 0)
    finally:
        t.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_abi_with_a_schedule(gpu, graphs):
    """Three moduli of the 2^14 class in one batch, the schedule changed
    between groups of squarings (graphs captured before a change are
    dropped); after each group every slot holds 3^(its bits) mod N."""
    rng = random.Random(47)
    a = 3
    Ns, plan = [], None
    seed = 100
    while len(Ns) < 3:
        N = usable_N(1000000, seed, a)
        seed += 1
        p = annie_amd.select_plan(N, a)
        if plan is None:
            plan = p
        if p == plan:
            Ns.append(N)
    K, pck, LN, lbase = plan
    assert LN == 14
    lib = annie_amd.load_library()
    h = ctypes.c_void_p()
    assert lib.annie_hip_batch_create(K, pck, LN, lbase, 3, 0, ctypes.byref(h)) == 0, annie_amd._err(lib)
    u32 = ctypes.c_uint32
    try:
        for s, N in enumerate(Ns):
            arr, n = annie_amd.int_to_limbs(N)
            f = (ctypes.c_uint64 * (n + 1))()
            nf = ctypes.c_size_t()
            assert lib.annie_hip_batch_load(h, s, arr, n, a, f, n + 1, ctypes.byref(nf)) == 0, annie_amd._err(lib)
        if graphs:
            assert lib.annie_hip_batch_set_graphs(h, 1) == 0
        ids = (u32 * 3)(0, 1, 2)
        assert lib.annie_hip_batch_set1(h, ids, 3) == 0
        stride = 1000000 // 64 + 4
        out = (ctypes.c_uint64 * (3 * stride))()
        nout = (ctypes.c_size_t * 3)()
        es = [0, 0, 0]  # set1: r0 = 3^0
        scheds = annie_amd.schedules(LN)
        for g, sch in enumerate([scheds[1], (0, 0), scheds[-1]]):
            assert lib.annie_hip_set_schedule(h, *sch) == 0, annie_amd._err(lib)
            for _ in range(6):
                bits = [rng.getrandbits(1) for _ in range(3)]
                m = (u32 * 3)(*[a if b else 1 for b in bits])
                assert lib.annie_hip_batch_square_mul(h, ids, m, 3) == 0, annie_amd._err(lib)
                es = [2 * e + b for e, b in zip(es, bits)]
            assert lib.annie_hip_batch_get(h, 0, ids, 3, out, stride, nout) == 0, annie_amd._err(lib)
            for s in range(3):
                v = sum(int(x) << (64 * i) for i, x in enumerate(out[s * stride:s * stride + nout[s]]))
                assert v == pow(a, es[s], Ns[s]), (g, sch, s)
            print(f"case graphs={graphs} schedule {sch}: 3 slots equal pow", flush=True)
    finally:
        lib.annie_hip_destroy(h)
