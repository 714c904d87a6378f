"""CPU tests of the product library: exports, generators, argument checking (no GPU compute)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in ("p2r.h", "p2r_gen.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(p2r_[a-z_A-Z0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from p2r_amd import _capi
    lib = C.CDLL(_capi.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_constants_match_the_header():
    """The Python mirror's copies of the header's enum constants carry the header's values."""
    import p2r_amd
    text = open(os.path.join(ROOT, "include", "p2r.h")).read()
    for name in ("P2R_MAX_NLAYER", "P2R_RAGGED_PAIR_STREAM", "P2R_RAGGED_LAYER_ORDER", "P2R_OPT_CHI2_PER_LAYER",
                 "P2R_GRAD_MIXED", "P2R_ERR_SHAPE"):
        m = re.search(name + r"\s*=\s*(-?\d+)", text)
        assert m and int(m.group(1)) == getattr(p2r_amd, name), name


def test_library_is_gfx950_code():
    from p2r_amd import _capi
    blob = open(_capi.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_reference_generator_matches_oracle(orc):
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(2, 64, 20, 32)
    trk, hits = p2r_amd.gen_reference(shape)
    t2, h2 = orc.gen_c0(2, 64, 20, 32)
    assert np.array_equal(trk.view(np.uint32), t2.view(np.uint32))
    assert np.array_equal(hits.view(np.uint32), h2.view(np.uint32))
    for lay in (0, 19):
        assert np.array_equal(p2r_amd.reference_hit(lay), orc.c0_hit(lay))


def test_uniform_generator_is_counter_based():
    import p2r_amd
    from p2r_amd import Shape
    a_t, a_h = p2r_amd.gen_uniform(Shape(4, 256, 20, 32))
    b_t, b_h = p2r_amd.gen_uniform(Shape(2, 256, 20, 32))
    # the first two events regenerate identically from a smaller request
    assert np.array_equal(a_t[:b_t.size].view(np.uint32), b_t.view(np.uint32))
    assert np.array_equal(a_h[:b_h.size], b_h)
    c_t, _ = p2r_amd.gen_uniform(Shape(2, 256, 20, 32), p2r_amd.gen_config(seed=7))
    assert not np.array_equal(c_t.view(np.uint32), b_t.view(np.uint32))


def test_uniform_generator_literal_statistics():
    """Literal mode: every parameter, covariance entry and hit coordinate x (1 + 0.1 N(0,1))."""
    import p2r_amd
    from p2r_amd import Shape, tracks_to_arrays
    shape = Shape(4, 2048, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape, p2r_amd.gen_config(mode=1))
    par, cov, q = tracks_to_arrays(trk, 32)
    base = np.array([-12.806846618652344, -7.723824977874756, 38.13014221191406, 0.23732035065189902,
                     -2.613372802734375, 0.35594117641448975])
    rel = par / base - 1
    assert np.all(np.abs(rel.mean(axis=0)) < 0.01)
    assert np.all(np.abs(rel.std(axis=0) - 0.1) < 0.01)
    assert set(np.unique(q)) == {-1, 1} and abs(q.mean()) < 0.05
    h = p2r_amd.hits_block_view(hits, 20, 32)
    r = np.hypot(h[:, :, 0, :], h[:, :, 1, :])
    np.testing.assert_allclose(r.mean(axis=(0, 2)), 20 + 5 * np.arange(20), rtol=0.02)
    assert np.all(h[:, :, 3, :] == np.float32(1e-4)) and np.all(h[:, :, 4, :] == 0)


def test_uniform_generator_consistent_mode(orc):
    """Consistent mode: positive-definite input covariances, hits on each track's helix within the hit
    resolution of their cylinder, and a chained fit that stays finite for every track."""
    import p2r_amd
    from p2r_amd import Shape, tracks_to_arrays
    shape = Shape(2, 1024, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    par, cov, q = tracks_to_arrays(trk, 32)
    for c in cov[:200]:
        P = np.zeros((6, 6))
        k = 0
        for i in range(6):
            for j in range(i + 1):
                P[i, j] = P[j, i] = c[k]
                k += 1
        assert np.linalg.eigvalsh(P).min() > 0
    h = p2r_amd.hits_block_view(hits, 20, 32)
    r = np.hypot(h[:, :, 0, :], h[:, :, 1, :])
    assert np.abs(r - (20 + 5 * np.arange(20))[None, :, None]).max() < 0.06  # 6 sigma of 0.01 cm
    p64, c64, chi2 = orc.run_f64(trk, hits, shape.nblk, 20, 32, semantics=1)
    assert np.isfinite(p64).all() and np.isfinite(chi2).all()
    assert np.median(chi2) / 20 < 20


def test_unsmeared_hits_lie_on_the_propagated_base_track(orc):
    """Generator hits (smear 0) = the oracle's FP64 propagation of the base track to each cylinder."""
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(1, 32, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape, p2r_amd.gen_config(mode=1, smear=0.0))
    par0 = trk.reshape(28, 32)[:6, 0].astype(np.float64)
    cov0 = trk.reshape(28, 32)[6:27, 0].astype(np.float64)
    h = p2r_amd.hits_block_view(hits, 20, 32)[0]
    for l in (0, 5, 19):
        hp = h[l, :3, 0].astype(np.float64)
        assert abs(np.hypot(hp[0], hp[1]) - (20 + 5 * l)) < 1e-4
        p, _, _ = orc.propagate_lane_f64(par0, cov0, 1, hp, float(orc.kfact_tbb()))
        # the reference's 5 fixed-point steps with a Taylor sin/cos land within
        # 1e-4 (relative) of the exact helix crossing the generator computes
        np.testing.assert_allclose(p[:3], hp, rtol=0, atol=1e-4 * np.linalg.norm(hp))


def test_ragged_generator_layout_and_masks():
    import p2r_amd
    counts = p2r_amd.lognormal_counts(50, median=300, sigma=0.5, lo=20, hi=2000)
    assert counts.min() >= 20 and counts.max() <= 2000
    trk, hits, mask, offs = p2r_amd.gen_ragged(counts, 20, 32, kmin=5, kmax=20)
    assert offs[0] == 0 and np.all(np.diff(offs) == (counts + 31) // 32)
    nhit = np.array([bin(int(m)).count("1") for m in mask])
    for e in range(len(counts)):
        s = nhit[offs[e] * 32: offs[e + 1] * 32]
        assert np.all(s[:counts[e]] >= 5) and np.all(s[:counts[e]] <= 20) and np.all(s[counts[e]:] == 0)
    real = nhit[nhit > 0]
    assert abs(real.mean() - 12.5) < 0.5


def test_create_without_device_returns_error():
    """No GPU in this container: p2r_create must fail with a status, not abort."""
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import p2r_amd
    with pytest.raises(p2r_amd.P2RError):
        p2r_amd.Context(0)
    lib = p2r_amd.lib()
    assert lib.p2r_forward(None, None, None, None, None, 0, None, None, None) == p2r_amd.P2R_ERR_ARG
    assert lib.p2r_error_string(2) == b"unsupported shape"


def test_torch_extension_loads_and_exports():
    """The thin torch extension is built in-tree and exposes its three functions (no GPU call)."""
    import os

    import p2r_amd.autograd as ag
    assert os.path.dirname(ag._ext.__file__).endswith(os.path.join("p2r-tests_amd", "p2r_amd"))
    for name in ("forward", "forward_adjoint", "default_kfact"):
        assert hasattr(ag._ext, name)
    import p2r_amd
    assert ag._ext.default_kfact() == p2r_amd.default_params().kfact


def test_python_buffer_checks_reject_host_memory_and_short_buffers():
    """The Python mirror refuses what would reach a kernel as a bad pointer: host memory where device memory is
    expected, and the shape rules it mirrors from the C-ABI decide when sizes are checked at all."""
    torch = pytest.importorskip("torch")
    import numpy as np
    import p2r_amd
    from p2r_amd import Shape, _check_dev, _check_host, _ragged_ok, _uniform_ok
    with pytest.raises(TypeError):
        _check_dev(torch.zeros(10), 5, "f32", 0, "trk")            # a CPU tensor as a device buffer
    with pytest.raises(TypeError):
        _check_dev(np.zeros(10, np.float32), 5, "f32", 0, "trk")   # a numpy array as a device buffer
    _check_dev(None, 5, "f32", 0, "chi2")                           # optional buffers may be absent
    _check_host(np.zeros(10, np.float32), 10, "host_trk")
    with pytest.raises(ValueError):
        _check_host(np.zeros(9, np.float32), 10, "host_trk")
    with pytest.raises(TypeError):
        _check_host(np.zeros(10, np.float64), 10, "host_trk")
    ok = lambda *a: _uniform_ok(p2r_amd.p2r_shape(*a))
    assert ok(1, 64, 20, 32) and ok(0, 32, 20, 32)
    assert not ok(1, 65, 20, 32) and not ok(1, 64, 0, 32) and not ok(1, 64, 20, 2048)
    assert not ok(70000, 65536, 20, 32) and not ok(1, 64, p2r_amd.P2R_MAX_NLAYER + 1, 32) and ok(1, 64, 64, 32)
    offs = np.array([0, 1, 3], np.int64)
    assert _ragged_ok(offs, 20, 32) and not _ragged_ok(offs, 33, 32) and not _ragged_ok(np.array([1, 2]), 20, 32)
    assert not _ragged_ok(np.array([0, 2, 1]), 20, 32)
