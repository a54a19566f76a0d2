"""The size selection (K, pck, LN, lbase) of the oracle and of the engine
against the reference's own create_transformGPU (src/transformGPU.h:822-885),
built from the reference sources into oracle/_ref/ref_plan.  The reference
prints its candidate per K and the chosen K_min before it constructs the
OpenCL engine; with no OpenCL platform (an empty ICD vendor directory) that
construction throws and the process ends normally.  Skipped where the
reference sources are absent."""
import os
import random
import re
import subprocess

import pytest

import annie_amd
import oracle

REF_PLAN = os.path.join(os.path.dirname(oracle.__file__), "_ref", "ref_plan")
pytestmark = pytest.mark.skipif(not os.path.exists(REF_PLAN), reason="reference sources not built (oracle/_ref)")


def ref_plan(N, a, icd_dir):
    env = dict(os.environ, OCL_ICD_VENDORS=str(icd_dir))
    out = subprocess.run([REF_PLAN, str(a)], input=format(N, "x"), capture_output=True, text=True, timeout=60,
                         env=env).stdout
    K = int(re.search(r"K_min = (\d+)", out).group(1))
    m = re.search(rf"K = {K}: ln = (\d+), lbase = (\d+), s = (\d+)", out)
    ln, lbase, s = (int(x) for x in m.groups())
    return K, s, ln, lbase


def _sizes():
    # every class boundary of the base-3 table up to 2^23 (first size of
    # each new plan, found by bisection on the oracle) and the sizes on
    # either side, plus random sizes
    rng = random.Random(7)
    out = set()
    bits, prev = 9000, None
    while bits < 170_000_000:
        p = oracle.plan((1 << (bits - 1)) | 1, 3)
        if prev is not None and p != prev:
            lo, hi = bits // 2, bits
            while hi - lo > 1:
                mid = (lo + hi) // 2
                if oracle.plan((1 << (mid - 1)) | 1, 3) == prev:
                    lo = mid
                else:
                    hi = mid
            out.update({lo, hi, hi + 1})
        prev = p
        bits = int(bits * 1.08) + 1
    out.update(rng.randrange(10_000, 160_000_000) for _ in range(20))
    return sorted(out)


def test_plan_matches_reference_create_transformGPU(tmp_path):
    rng = random.Random(8)
    checked = 0
    for bits in _sizes():
        N = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        want = ref_plan(N, 3, tmp_path)
        assert oracle.plan(N, 3) == want, (bits, want)
        assert annie_amd.select_plan(N, 3) == want, (bits, want)
        checked += 1
    assert checked > 80


@pytest.mark.parametrize("a", [2, 5, 7, 11])
def test_plan_matches_reference_other_bases(tmp_path, a):
    rng = random.Random(a)
    for _ in range(12):
        bits = rng.randrange(20_000, 40_000_000)
        N = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        want = ref_plan(N, a, tmp_path)
        assert oracle.plan(N, a) == want and annie_amd.select_plan(N, a) == want, (bits, a, want)
