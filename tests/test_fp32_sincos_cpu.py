"""The fast FP32 instance's sin/cos (p2r_lane_math.inc, the fast branch) on the CPU.

The fast instance (p2r_forward's P2R_FAST, and the per-layer values of every adjoint's reverse sweep) computes
sin/cos with its own three-part Cody-Waite reduction and Taylor polynomials instead of the library's sincosf.
tests/fp32_sincos_harness.cpp builds that code on the host and compares it with libm's double sin/cos over
[-4, 4], [-100, 100], [-6000, 6000] and the floats nearest to multiples of pi/2: every result within 2 units
of 2^-24 (one ulp of a value near 1).  A wrong coefficient, reduction constant or quadrant would be off by far
more.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fast_sincos_matches_libm(tmp_path):
    exe = str(tmp_path / "fp32sc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, os.path.join(HERE, "fp32_sincos_harness.cpp")],
                   check=True)
    out = subprocess.run([exe, "200000"], capture_output=True, text=True, check=True).stdout.split()
    e_sin, e_cos, nan = float(out[0]), float(out[1]), int(out[2])
    print(f"max |error| / 2^-24: sin {e_sin}, cos {e_cos}", flush=True)
    assert e_sin <= 2.0 and e_cos <= 2.0
    assert nan == 1, "NaN and infinite arguments give NaN"
