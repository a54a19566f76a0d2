"""Kernel-level parity: each stage of one MMGW squaring on the GPU against the
CPU oracle's same intermediate, word for word (convolution residues, T_R
residues, T_Q digits, the m' convolution, m' + 2)."""
import pytest

import annie_amd
import oracle
from tests.helpers import assert_same_words, usable_N

pytestmark = pytest.mark.gpu

CASES = [(20000, 3), (60000, 4), (150000, 5), (300000, 6), (1000000, 8), (4000000, 10)]


@pytest.mark.parametrize("bits,seed", CASES)
def test_square_stages(gpu, bits, seed):
    a = 3
    N = usable_N(bits, seed, a)
    t = annie_amd.Transform(N, a, min_ln=5)
    o = oracle.NttTransform(N, a)
    print(f"case bits={bits} K={t.K} LN={t.LN} lbase={t.lbase}", flush=True)
    for m in (t, o):
        m.set1()  # identical starting register on both sides
    n, K = t.n, t.K
    for stage in (1, 2, 3, 4):
        t.debug_square_stage(a, stage)
        out_r, out_q = o.debug_square_stage(a, stage)
        if stage == 1:
            assert_same_words(t.read_register(5), out_r, "cyclic convolution", n)
            assert_same_words(t.read_register(6), list(out_q)[:K * n], "negacyclic convolution", n)
        elif stage == 2:
            assert_same_words(t.read_register(4), out_r, "T_R", n)
            assert_same_words(t.read_tq(), list(out_q)[:3 * n], "T_Q digits", n)
        elif stage == 3:
            assert_same_words(t.read_register(5), out_r, "m' convolution", n)
        else:
            assert_same_words(t.read_register(4), out_r, "m' + 2", n)
