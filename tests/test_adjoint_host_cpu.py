"""The hand-written reverse sweep (p2r_adjoint.inc) checked on the CPU, before any GPU run.

tests/adjoint_host_harness.cpp builds the product's per-lane bodies with real = double, as the library's FP64
instance does, and computes the gradient of each track's total chi2 with layer_forward_values forward and
layer_adjoint backward.  The oracle's forward-mode Dual<16> (FP64, oracle/p2r_oracle.hpp) differentiates its own
restatement of the same algorithm; the two agree to FP64 rounding on every track when the reverse is right, and a
dropped or mis-signed term of the reverse shows up as an O(1) relative error in some component.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("adjh") / "adjh")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, os.path.join(HERE, "adjoint_host_harness.cpp")], check=True)
    return exe


def _run(exe, T, H, kf, tracks, nlayer):
    bs = T.shape[2]
    lines = [str(len(tracks))]
    for t in tracks:
        b, i = divmod(int(t), bs)
        lines.append(" ".join(repr(float(x)) for x in T[b, :27, i]) + f" {int(T[b, 27, i].view(np.int32))} {kf!r} "
                     f"{nlayer}")
        lines.extend(" ".join(repr(float(x)) for x in H[b, l, :, i]) for l in range(nlayer))
    out = subprocess.run([exe], input="\n".join(lines), capture_output=True, text=True, check=True).stdout
    return np.array([[float(v) for v in l.split()] for l in out.splitlines() if l.strip()])


@pytest.mark.parametrize("nlayer", [1, 3, 20])
def test_reverse_sweep_matches_dual(harness, orc, nlayer):
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(1, 4096, nlayer, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    T, H = trk.reshape(-1, 28, 32), hits.reshape(-1, nlayer, 9, 32)
    kf = float(p2r_amd.default_params().kfact)
    tracks = np.sort(np.random.default_rng(0x484F5354 + nlayer).choice(shape.ntrk_total, 48, replace=False))
    g = _run(harness, T, H, kf, tracks, nlayer)
    errs, worst = [], 0.0
    for row, t in zip(g, tracks):
        b, i = divmod(int(t), 32)
        ref, _ = orc.chi2_grad_dual(T[b, :6, i], T[b, 6:27, i], int(T[b, 27, i].view(np.int32)),
                                    np.ascontiguousarray(H[b, :, :, i]), nlayer, kf)
        errs.append(np.linalg.norm(row - ref) / np.linalg.norm(ref))
        # every component, against the gradient's scale: a wrong term would stand out in its own component
        worst = max(worst, float(np.max(np.abs(row - ref)) / np.max(np.abs(ref))))
    errs = np.array(errs)
    print(f"nlayer {nlayer}: normwise vs Dual median {np.median(errs):.2e} max {errs.max():.2e}; "
          f"worst component {worst:.2e}", flush=True)
    assert errs.max() < 1e-6 and worst < 1e-6


def test_a_dropped_term_is_seen(harness, orc, tmp_path):
    """The comparison's power: the same harness built with one term of the reverse of the iterations taken out
    (the r0 -> (xo, yo) chain of the converged-iteration block) disagrees with Dual."""
    src = open(os.path.join(HERE, "adjoint_host_harness.cpp")).read()
    adj = open(os.path.join(HERE, "..", "p2r-tests_amd", "csrc", "p2r_adjoint.inc")).read()
    marker = "xob += r0b * xo * it.rinv;"
    assert adj.count(marker) == 1
    (tmp_path / "p2r_adjoint_planted.inc").write_text(adj.replace(marker, "/* planted: term dropped */"))
    planted = src.replace('#include "../p2r-tests_amd/csrc/p2r_adjoint.inc"', '#include "p2r_adjoint_planted.inc"')
    planted = planted.replace('#include "../p2r-tests_amd/csrc/', '#include "' +
                              os.path.join(HERE, "..", "p2r-tests_amd", "csrc") + "/")
    (tmp_path / "h.cpp").write_text(planted)
    exe = str(tmp_path / "h")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", str(tmp_path), "-o", exe, str(tmp_path / "h.cpp")], check=True)
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(1, 1024, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    T, H = trk.reshape(-1, 28, 32), hits.reshape(-1, 20, 9, 32)
    kf = float(p2r_amd.default_params().kfact)
    tracks = np.arange(0, 1024, 64)
    g = _run(exe, T, H, kf, tracks, 20)
    errs = []
    for row, t in zip(g, tracks):
        b, i = divmod(int(t), 32)
        ref, _ = orc.chi2_grad_dual(T[b, :6, i], T[b, 6:27, i], int(T[b, 27, i].view(np.int32)),
                                    np.ascontiguousarray(H[b, :, :, i]), 20, kf)
        errs.append(np.linalg.norm(row - ref) / np.linalg.norm(ref))
    assert max(errs) > 1e-4
