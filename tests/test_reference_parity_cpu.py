"""The parity bar against the reference's own CPU path, measured on the CPU at the benchmark's sizes
(10 x 9600 x 20 and 100 x 9600 x 20, seeded benchmark generator, both layer semantics).

The chain behind it:
  * the GPU's P2R_EXACT output is word-for-word the oracle's IEEE-strict FP32 run with correctly rounded
    sin/cos (tests/test_gpu_parity.py, at these same sizes);
  * the oracle's IEEE-strict run with libm's sinf/cosf is word-for-word the reference's own compiled
    propagateToR / KalmanUpdate (tests/test_oracle_pins.py, against propagate-tor-test_acc.v3.cpp built
    where it lies, whose routines are tbb.cpp:296-812's);
so comparing the oracle's two trig modes here compares the GPU with the reference's CPU path.  The two
runs differ only in the four sin/cos values of each layer (tbb.cpp:689, 768-769), which libm rounds
differently from the correctly rounded value on about 0.1 % of arguments.

What is measured, per element, is |gpu - ref| / s with s the element's natural scale (parity.scales:
|value| for a parameter, sqrt(|P_ii P_jj|) for a covariance entry, floored at 1e-6 of the field's median).
The bar is 1e-4.  The yardstick for what FP32 arithmetic can promise is the reference's own CPU path as
its build compiles it (-O3 -ffast-math, propagate-tor-test build.py:91; here the oracle built with those
flags, liboracle_bench.so): against its own IEEE-strict build it keeps far fewer elements within 1e-4.
The covariance update P - K H P (tbb.cpp:530-619) cancels for precise hits, so a one-ulp change of a
sine moves such covariance entries by far more than 1e-4 of their scale in any FP32 build; the
parameters stay within the bar on all but about one element in 100 000, and the summary lines print the
same digits.
"""
import numpy as np
import pytest

import oracle
import parity

RTOL = 1e-4


@pytest.fixture(scope="module")
def orc_alt():
    return oracle.Oracle("liboracle_bench.so")


def parity_vs_reference(orc, orc_alt, trk, hits, nblk, nlayer, bs, semantics):
    """Numbers of the GPU (exact mode) against the reference's CPU path, and of the reference's own
    fast-math build against the same strict build (the yardstick), on one batch."""
    ref = orc.run_f32(trk, hits, nblk, nlayer, bs, semantics=semantics)                # reference, libm trig
    gpu = orc.run_f32(trk, hits, nblk, nlayer, bs, semantics=semantics, cr_trig=True)  # = GPU exact, word for word
    fm = orc_alt.run_f32(trk, hits, nblk, nlayer, bs, semantics=semantics)             # reference flags
    par64, cov64, _ = orc.run_f64(trk, hits, nblk, nlayer, bs, semantics=semantics)
    r = parity.vs_reference(gpu, ref, fm, bs)
    r["reference_fastmath_build"] = r.pop("reference_build_flags")
    fin = lambda b: np.isfinite(np.hstack(parity._arrays(b, bs)[:2])).all(axis=1)
    r.update({
        "nan_pattern_equal": bool(np.array_equal(fin(gpu), fin(ref))),
        "q_equal": bool(np.array_equal(parity._arrays(gpu, bs)[2], parity._arrays(ref, bs)[2])),
        "err_vs_fp64_gpu": parity.errors_vs_fp64(gpu, bs, par64, cov64),
        "err_vs_fp64_ref": parity.errors_vs_fp64(ref, bs, par64, cov64),
        "summary_gpu": _summary(orc, gpu, nblk, bs, hits),
        "summary_ref": _summary(orc, ref, nblk, bs, hits),
    })
    return r


def _summary(orc, out, nblk, bs, hits):
    """The reference's printed summary lines (tbb.cpp:1036-1048) plus the RMS of pT, phi and theta, in the
    same %f format; the reference hit is the batch's first hit."""
    st, cn = orc.summary(out, 1, nblk * bs, bs, hits[:3 * bs:bs])
    return oracle.summary_lines(st, cn) + ["track pt rms=%f phi rms=%f theta rms=%f" % tuple(st[19:22])]


@pytest.mark.slow
@pytest.mark.parametrize("semantics", [0, 1], ids=["reference", "chained"])
@pytest.mark.parametrize("nevts", [10, 100])
def test_exact_path_within_1e4_of_reference_cpu(orc, orc_alt, nevts, semantics):
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(nevts, 9600, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    r = parity_vs_reference(orc, orc_alt, trk, hits, shape.nblk, shape.nlayer, shape.bsize, semantics)
    fm = r["reference_fastmath_build"]
    print(f"{nevts}x9600x20 sem={semantics}: params within 1e-4 {r['params_frac_within_1e-4']:.7f} "
          f"(max {r['params_max_rel']:.2e}), cov {r['cov_frac_within_1e-4']:.4f}, worst {r['worst_element']}; "
          f"reference's fast-math build: params {fm['params_frac_within_1e-4']:.7f} (max {fm['params_max_rel']:.2e}), "
          f"cov {fm['cov_frac_within_1e-4']:.4f}", flush=True)
    assert r["nan_pattern_equal"] and r["q_equal"]
    # the summary lines (means and spreads of x, y, z, r, pT, phi, theta) print the same digits
    assert r["summary_gpu"] == r["summary_ref"]
    # parameters: within the bar on all but about one element in 100 000 (near-singular fits, where a
    # one-ulp change moves the result), and never further off than the fast-math build
    assert r["params_frac_within_1e-4"] >= 0.99998
    assert r["params_max_rel"] <= fm["params_max_rel"]
    # every field at least as close to the reference's strict build as its own fast-math build is
    assert np.all(np.array(r["per_field_frac_within_1e-4"]) >= np.array(fm["per_field_frac_within_1e-4"]))
    assert r["cov_frac_within_1e-4"] >= 0.97
    # and as accurate against the FP64 replay as the reference's CPU path (the RMS of a few chaotic
    # tracks moves by some percent with a one-ulp change of a sine: 25 % of slack)
    for q in ("x", "y", "z", "ipt", "phi", "theta", "cov"):
        for st in ("p99", "rms"):
            assert r["err_vs_fp64_gpu"][q][st] <= 1.25 * r["err_vs_fp64_ref"][q][st] + 1e-9, (q, st)


def test_planted_mistake_fails_the_bar(orc, orc_alt):
    """The comparison above can fail: a result with one covariance term perturbed by a relative 1e-3 on
    every track (a dropped or wrong term would do more) loses the per-field comparison."""
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(1, 4096, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    ref = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=0)
    fm = orc_alt.run_f32(trk, hits, shape.nblk, 20, 32, semantics=0)
    bad = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=0, cr_trig=True).reshape(-1, 28, 32).copy()
    bad[:, 6 + 14] *= np.float32(1 + 1e-3)  # covariance (4, 4): phi-phi
    frac_bad = (parity.rel_to(bad.ravel(), ref, 32) <= RTOL).mean(axis=0)
    frac_fm = (parity.rel_to(fm, ref, 32) <= RTOL).mean(axis=0)
    assert frac_bad[6 + 14] < frac_fm[6 + 14]
