"""The reference's own known answers (the QUICK_TEST list of
src/main.cpp:163) through the GPU path: F(n) for n in the list is a
(probable) prime except F(802) and F(1601); every Gerbicz-Li check passes.
Entries up to F(397379) (275,000 bits) run here, once through the plain
engine (every size on the GPU) and once all together through the batch
driver."""
import pytest

import annie_amd

pytestmark = pytest.mark.gpu

QUICK = [359, 431, 802, 1601, 2971, 5387, 9311, 25561, 35999, 81839, 148091, 397379]
COMPOSITE = {802, 1601}


def _fib(n):
    a, b = 0, 1
    for _ in range(n):
        a, b = b, a + b
    return a


@pytest.mark.parametrize("i", QUICK)
def test_quick_test_list_plain(gpu, i):
    r = annie_amd.check(_fib(i), gmp_threshold_digits=1, min_ln=4)
    print(f"case fibonacci({i}): {r['message'][-70:]}", flush=True)
    assert r["status"] == "success" and r["gl_ok"]
    assert r["is_prp"] == (i not in COMPOSITE)


def test_quick_test_list_batched(gpu):
    res, st = annie_amd.batch_run([{"N": _fib(i)} for i in QUICK], engine="gpu")
    for i, r in zip(QUICK, res):
        print(f"case fibonacci({i}): {r['message'][-70:]}", flush=True)
        assert r["status"] == "success" and r["gl_ok"], r
        assert r["is_prp"] == (i not in COMPOSITE)
