"""GPU tests of the C++ driver lib/p2r_bench (the reference program's command line and stdout protocol,
tbb.cpp:814-1055): on the reference's own C0 inputs (10 events x 9600 tracks x 20 layers) it prints the
`done ntracks=... tot time=...` line that build.py:83 parses and the validation summary of
tbb.cpp:1036-1048.  The summary must equal the oracle's summary of the same output (the exact path is
bit-identical to the oracle with correctly rounded sin/cos), and agree with the reference program's own
printed lines (tests/golden/c0_summary_ref.txt, made with glibc's sinf/cosf) to the last printed digit.
The kernel-only timing replayed from a captured graph (--graph) prints the same summary."""
import os
import re
import subprocess

import numpy as np
import pytest

from p2r_amd import Shape

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "p2r-tests_amd", "lib", "p2r_bench")
GOLDEN = os.path.join(ROOT, "tests", "golden", "c0_summary_ref.txt")


def _run(args, env=None):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(EXE), f"driver not built: {EXE}"
    p = subprocess.run([EXE] + args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, universal_newlines=True,
                       timeout=120, env=None if env is None else dict(os.environ, **env))
    assert p.returncode == 0, p.stdout
    return p.stdout


def _summary(out):
    lines = out.splitlines()
    i = next(k for k, l in enumerate(lines) if l.startswith("track x avg="))
    return lines[i:i + 13]


@pytest.fixture(scope="module")
def c0_expected():
    import oracle
    import p2r_amd
    orc = oracle.Oracle()
    shape = Shape(10, 9600, 20, 32)
    trk, hits = p2r_amd.gen_reference(shape)
    cpu = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=0, cr_trig=True)
    st, cn = orc.summary(cpu, shape.nevts, shape.ntrks, shape.bsize, orc.c0_hit(19))
    return oracle.summary_lines(st, cn)


def _numbers(line):
    return [float(x) for x in re.findall(r"[-+]?\d+\.\d+|[-+]?\d+", line)]


@pytest.mark.parametrize("graph", [False, True])
def test_driver_c0_protocol_and_summary(c0_expected, graph):
    out = _run(["--input", "reference", "--nevts", "10", "--NITER", "2"] + (["--graph"] if graph else []))
    from p2r_amd import scan
    m = scan.RESULT_RE.search(out)
    assert m and int(m.group("ntracks")) == 10 * 9600 * 2 and float(m.group("time")) > 0
    assert re.search(r"^formatted 2 10 9600 32 300 ", out, re.M)
    got = _summary(out)
    assert got == c0_expected, "\n".join(got)
    want = open(GOLDEN).read().split("\n")[:13]
    for g, w in zip(got, want):
        assert re.sub(r"[-\d.]+", "#", g) == re.sub(r"[-\d.]+", "#", w)  # same text around the numbers
        np.testing.assert_allclose(_numbers(g), _numbers(w), rtol=0, atol=2e-6)


def test_driver_baseline_inputs_and_transfers():
    """Seeded benchmark inputs, every timed-region choice of the reference (transfers none / h2d / d2h / both with
    streams): each run prints a done line and the same summary (the output does not depend on the timing)."""
    outs = [_run(["--input", "baseline", "--nevts", "4", "--NITER", "1"] + extra)
            for extra in ([], ["--transfers", "h2d"], ["--transfers", "d2h"],
                          ["--transfers", "both", "--num-streams", "3"], ["--graph"])]
    sums = [_summary(o) for o in outs]
    assert all(s == sums[0] for s in sums[1:])
    assert all("done ntracks=38400 " in o for o in outs)


def test_driver_with_roctx_ranges_enabled():
    """P2R_ROCTX=1 opens the marker library and wraps every entry point in a range; outside a profiler the
    ranges are no-ops and the run prints the same summary."""
    args = ["--input", "baseline", "--nevts", "2", "--NITER", "1"]
    assert _summary(_run(args, {"P2R_ROCTX": "1"})) == _summary(_run(args))
