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


# The rest of the QUICK_TEST list (src/main.cpp:163) above 400,000 bits, run
# through the command-line front end (the expression evaluated by the CLI's
# parser, the test on the GPU engine, src/main.cpp:246-264): each is a
# Fermat probable prime with every Gerbicz-Li check passing (the two largest,
# about 1 and 2.5 minutes, with the long tests).
CLI_QUICK = [590041, pytest.param(1285607, marks=pytest.mark.long), pytest.param(2904353, marks=pytest.mark.long)]


@pytest.mark.parametrize("i", CLI_QUICK)
def test_quick_test_list_cli(gpu, capsys, tmp_path, i):
    import os
    import subprocess
    exe = os.path.join(annie_amd.LIBDIR, "annie")
    expr = f"fibonacci({i})"
    lines = []
    with capsys.disabled():
        print(f"case CLI {expr}", flush=True)
        p = subprocess.Popen([exe, "-q", expr, "-w", str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True)
        for ln in p.stdout:  # progress lines every 10 s keep the log moving
            print("  " + ln.rstrip(), flush=True)
            lines.append(ln)
        rc = p.wait(timeout=600)
    out = "".join(lines)
    assert rc == 0, out[-2000:]
    assert f"{expr} is a Fermat probable prime, a = 2." in out, out[-2000:]
    assert "validation failed" not in out.lower()
