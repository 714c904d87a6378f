"""Operation counts of the product's own per-lane code (p2r_lane_math.inc, p2r_adjoint.inc), frozen.

tests/opcount_harness.cpp instantiates those bodies on the host with a counting scalar type and runs the
chained fit of tracks from the benchmark generator layer by layer.  Per track-layer it reports the
operations of the forward update, of the layer's forward values (layer_forward_values), of layer_adjoint,
how many fixed-point iterations had converged (the reverse of an iteration's derivative terms runs only
then: |r - r0| < 1e-4, tbb.cpp:714), and the parts that layer_adjoint recomputes or leaves out: each
iteration's values recomputed from its start point inside the reverse of the iterations
(iteration_values), and the updated state, which the reverse does not need (update_out).

Frozen here and used by bench.py for the fwd+adjoint roofline:
  forward update               1214  (the reference algorithm's 1212 useful FLOPs, oracle-counted, differ
                                      by the chi2 the product adds and by structural terms it skips)
  forward values of a layer    1203  (192 of them the updated state)
  recomputed in the reverse    84    (each iteration from its start point)
  reverse sweep of a layer     1878 + 164 per converged iteration
  mean reverse on the benchmark's tracks (1024 tracks x 20 layers, chained): 2107 (1.40 converged iterations
  per layer)
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REVERSE_BASE, REVERSE_PER_CONVERGED, REVERSE_MEAN = 1878, 164, 2107
VALUES, ITERATIONS, UPDATE_OUT = 1203, 84, 192


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("opcount") / "opcount")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, os.path.join(HERE, "opcount_harness.cpp")], check=True)
    return exe


def _counts(exe, ntrk, nlayer=20):
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(1, ntrk, nlayer, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    T = trk.reshape(-1, 28, 32)
    H = hits.reshape(-1, nlayer, 9, 32)
    kf = p2r_amd.default_params().kfact
    lines = [str(ntrk)]
    for t in range(ntrk):
        b, i = divmod(t, 32)
        lines.append(" ".join(repr(float(x)) for x in T[b, :27, i]) + f" {int(T[b, 27, i].view(np.int32))} {kf!r} "
                     f"{nlayer}")
        lines.extend(" ".join(repr(float(x)) for x in H[b, l, :, i]) for l in range(nlayer))
    out = subprocess.run([exe], input="\n".join(lines), capture_output=True, text=True, check=True).stdout
    return np.array([[int(v) for v in l.split()] for l in out.splitlines() if l.strip()])


def _reverse(a):
    """Operations of the reverse proper: layer_adjoint minus the forward values it computes (the layer's
    values without the updated state, plus each iteration recomputed from its start point)."""
    fwd, values, adjoint, cond, iters, upd_out = a.T
    return adjoint - (values - upd_out) - iters


def test_adjoint_opcount_frozen(harness):
    a = _counts(harness, 1024)
    fwd, values, adjoint, cond, iters, upd_out = a.T
    assert len(a) == 1024 * 20
    assert set(fwd) == {1214} and set(values) == {VALUES}
    assert set(iters) == {ITERATIONS} and set(upd_out) == {UPDATE_OUT}
    reverse = _reverse(a)
    assert np.array_equal(reverse, REVERSE_BASE + REVERSE_PER_CONVERGED * cond)
    mean = float(reverse.mean())
    print(f"reverse ops per layer: mean {mean:.1f}, converged iterations per layer {cond.mean():.3f}, "
          f"distribution {np.bincount(cond).tolist()}", flush=True)
    assert abs(mean - REVERSE_MEAN) <= 1


def test_counter_sees_a_changed_formula(harness):
    """The harness counts what the bodies execute: the reverse without converged iterations is exactly the
    frozen base, and each converged iteration adds the same block (so a dropped or extra term changes a
    count above)."""
    a = _counts(harness, 64)
    rev = _reverse(a)
    for c in np.unique(a[:, 3]):
        assert set(rev[a[:, 3] == c]) == {REVERSE_BASE + REVERSE_PER_CONVERGED * int(c)}


def test_bench_uses_the_frozen_counts():
    import bench
    assert bench.REVERSE_FLOP_PER_UPDATE == REVERSE_MEAN and bench.FLOP_PER_UPDATE == 1212
