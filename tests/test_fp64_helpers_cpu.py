"""The FP64 instance's division, reciprocal, square root and sin/cos (p2r_lane_math.inc, P2R_DOUBLE_EST) on the CPU.

The library's FP64 instance, which computes the F64 and MIXED gradients, refines the hardware estimates of 1/x and
1/sqrt(x) by Newton steps instead of calling the IEEE library sequences.  tests/fp64_helpers_harness.cpp builds
those helpers on the host with the estimates emulated as the exact value times (1 + d), |d| <= 2^-E, and reports
their largest error in ulps of the IEEE result over random operands from 1e-30 to 1e30.  With E = 20 (the hardware
estimates are closer) every result is within 1 ulp.  With E = 4 the same steps cannot converge, and the errors run
to thousands of ulps: the check is not blind to a refinement that falls short.  The instance's own sin/cos
(Cody-Waite reduction, Taylor polynomials) is compared with libm's over [-1e5, 1e5], over [-4, 4] and at the
doubles nearest to multiples of pi/2: a wrong coefficient, reduction constant or quadrant would be off by far more
than the few units of 2^-52 allowed.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fp64h") / "fp64h")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, os.path.join(HERE, "fp64_helpers_harness.cpp")],
                   check=True)
    return exe


def _run(exe, n, bits, sincos=False):
    out = subprocess.run([exe, str(n), str(bits)], capture_output=True, text=True, check=True).stdout.split("\n")
    if sincos:
        e_sin, e_cos, nan = out[2].split()
        return float(e_sin), float(e_cos), int(nan)
    e_rcp, e_div, e_sqrt = (float(v) for v in out[0].split())
    specials = [int(v) for v in out[1].split()]
    return e_rcp, e_div, e_sqrt, specials


def test_refined_estimates_are_within_one_ulp(harness):
    e_rcp, e_div, e_sqrt, specials = _run(harness, 200000, 20)
    print(f"max ulps: rcp {e_rcp}, div {e_div}, sqrt {e_sqrt}", flush=True)
    assert e_rcp <= 1.0 and e_div <= 1.0 and e_sqrt <= 1.0
    assert specials == [1, 1, 1], "sqrt(0) = 0, sqrt(inf) = inf, sqrt(-1) = NaN"


def test_coarse_estimates_are_caught(harness):
    e_rcp, e_div, e_sqrt, _ = _run(harness, 20000, 4)
    assert min(e_rcp, e_div, e_sqrt) > 1000.0


def test_sincos_matches_libm(harness):
    e_sin, e_cos, nan = _run(harness, 200000, 20, sincos=True)
    print(f"max |error| / 2^-52: sin {e_sin}, cos {e_cos}", flush=True)
    assert e_sin <= 4.0 and e_cos <= 4.0
    assert nan == 1, "NaN and infinite arguments give NaN"

