"""Host-side argument checking of every C-ABI entry point: null pointers, bad shapes, out-of-range enums
and inconsistent layouts come back as status codes with a message (p2r_last_error), never an abort,
and launch nothing.  The context stays usable afterwards.  (The reference aborts through hipCheck,
include/hipCheck.h:20-28.)"""
import ctypes as C

import numpy as np
import pytest

from p2r_amd import Shape

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture()
def ctx():
    import p2r_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = p2r_amd.Context(0)
    yield c
    c.close()


def _status(fn):
    import p2r_amd
    try:
        fn()
    except p2r_amd.P2RError as e:
        assert "p2r status" in str(e)
        return e.status
    return 0


def test_forward_argument_errors(ctx):
    import p2r_amd
    from p2r_amd import _capi
    L = _capi.lib()
    t = torch.zeros(28 * 32 * 2, device="cuda")
    h = torch.zeros(9 * 32 * 20 * 2, device="cuda")
    sh = _capi.p2r_shape(1, 64, 20, 32)
    p = p2r_amd.default_params()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda x: C.c_void_p(x.data_ptr())
    assert L.p2r_forward(ctx._h, None, P(h), P(t), C.byref(sh), 0, C.byref(p), None, stream) == p2r_amd.P2R_ERR_ARG
    assert b"null pointer" in L.p2r_last_error(ctx._h)
    assert L.p2r_forward(ctx._h, P(t), P(h), P(t), None, 0, C.byref(p), None, stream) == p2r_amd.P2R_ERR_ARG
    assert L.p2r_forward(None, P(t), P(h), P(t), C.byref(sh), 0, C.byref(p), None, stream) == p2r_amd.P2R_ERR_ARG
    for bad in (Shape(1, 64, 0, 32), Shape(1, 64, 20, 0), Shape(1, 64, 20, 2048), Shape(-1, 64, 20, 32),
                Shape(1, 65, 20, 32), Shape(70000, 65536, 20, 32), Shape(1, 64, p2r_amd.P2R_MAX_NLAYER + 1, 32),
                Shape(1, 64, 60000, 32)):
        assert _status(lambda: ctx.forward(t, h, t, bad)) == p2r_amd.P2R_ERR_SHAPE, bad
    assert _status(lambda: ctx.forward(t, h, t, Shape(1, 64, 20, 32), semantics=5)) == p2r_amd.P2R_ERR_ARG
    for prm in (p2r_amd.default_params(precision=7), p2r_amd.default_params(b_scale=0.0),
                p2r_amd.default_params(options=8)):
        assert _status(lambda: ctx.forward(t, h, t, Shape(1, 64, 20, 32), params=prm)) == p2r_amd.P2R_ERR_ARG
        # an empty batch is no excuse for an invalid request
        assert _status(lambda: ctx.forward(t, h, t, Shape(0, 64, 20, 32), params=prm)) == p2r_amd.P2R_ERR_ARG
    ctx.forward(t, h, t, Shape(1, 64, 20, 32))  # still usable


def test_ragged_argument_errors(ctx):
    import p2r_amd
    t = torch.zeros(28 * 32 * 2, device="cuda")
    h = torch.zeros(9 * 32 * 20 * 2, device="cuda")
    m = torch.zeros(64, dtype=torch.int32, device="cuda")
    ok = np.array([0, 1, 2], np.int64)
    cases = [
        (dict(blk_offsets=np.array([1, 2], np.int64)), p2r_amd.P2R_ERR_SHAPE),   # offsets[0] != 0
        (dict(blk_offsets=np.array([0, 2, 1], np.int64)), p2r_amd.P2R_ERR_SHAPE),  # decreasing
        (dict(nlayer=33), p2r_amd.P2R_ERR_SHAPE),                                  # beyond the 32-bit mask
        (dict(nlayer=0), p2r_amd.P2R_ERR_SHAPE),
        (dict(bsize=0), p2r_amd.P2R_ERR_SHAPE),
        (dict(tile_rounds=3), p2r_amd.P2R_ERR_ARG),
        (dict(tile_rounds=-3), p2r_amd.P2R_ERR_ARG),
        (dict(semantics=9), p2r_amd.P2R_ERR_ARG),
        (dict(params=p2r_amd.default_params(options=16)), p2r_amd.P2R_ERR_ARG),
    ]
    for kw, want in cases:
        a = dict(blk_offsets=ok, nlayer=20, bsize=32, semantics="chained", tile_rounds=0, params=None)
        a.update(kw)
        got = _status(lambda: ctx.forward_ragged(t, h, t, a["blk_offsets"], m, a["nlayer"], a["bsize"],
                                                 a["semantics"], params=a["params"], tile_rounds=a["tile_rounds"]))
        assert got == want, (kw, got)
    assert _status(lambda: ctx.forward_ragged(t, h, t, ok, None, 20, 32, "chained")) == p2r_amd.P2R_ERR_ARG


def test_adjoint_and_graph_argument_errors(ctx):
    import p2r_amd
    t = torch.zeros(28 * 32 * 2, device="cuda")
    o = torch.zeros(28 * 32 * 2, device="cuda")
    h = torch.zeros(9 * 32 * 65 * 2, device="cuda")
    s20 = Shape(1, 64, 20, 32)
    assert _status(lambda: ctx.forward_adjoint(t, h, o, Shape(1, 64, 65, 32))) == p2r_amd.P2R_ERR_SHAPE
    assert _status(lambda: ctx.forward_adjoint(t, h, o, Shape(1, 65, 20, 32))) == p2r_amd.P2R_ERR_SHAPE
    assert _status(lambda: ctx.forward_adjoint(t, h, o, s20, grad_precision=5)) == p2r_amd.P2R_ERR_ARG
    assert _status(lambda: ctx.forward_adjoint(t, h, o, s20, checkpoint_stride=3)) == p2r_amd.P2R_ERR_ARG
    assert _status(lambda: ctx.forward_adjoint(t, h, o, s20, params=p2r_amd.default_params(options=2))) == \
        p2r_amd.P2R_ERR_UNSUPPORTED
    # the reverse sweep rereads the input tracks: an output on top of them is refused, whole or in part
    assert _status(lambda: ctx.forward_adjoint(t, h, t, s20)) == p2r_amd.P2R_ERR_ARG
    assert b"overlaps" in p2r_amd._capi.lib().p2r_last_error(ctx._h)
    big = torch.zeros(28 * 32 * 3, device="cuda")
    assert _status(lambda: ctx.forward_adjoint(big[:28 * 64], h, big[28 * 32:], s20)) == p2r_amd.P2R_ERR_ARG
    assert _status(lambda: ctx.graph_forward_adjoint(t, h, t, s20)) == p2r_amd.P2R_ERR_ARG
    assert _status(lambda: ctx.graph_forward_adjoint(t, h, o, s20, checkpoint_stride=5)) == p2r_amd.P2R_ERR_ARG
    assert _status(lambda: ctx.graph_forward_ragged(t, h, t, np.array([0, 1], np.int64), None, 20, 32)) == \
        p2r_amd.P2R_ERR_ARG
    from p2r_amd import _capi
    assert _capi.lib().p2r_graph_launch(None, None) == p2r_amd.P2R_ERR_ARG
    assert _capi.lib().p2r_graph_destroy(None) == p2r_amd.P2R_ERR_ARG


def test_tuning_streamed_and_repack_errors(ctx):
    import p2r_amd
    for bad in ({"fwd_threads": 32}, {"fwd_threads": 512}, {"ragged_threads": 96}, {"ragged_schedule": -3},
                {"checkpoint_stride": 16}, {"fwd_waves": 5}):
        assert _status(lambda: ctx.set_tuning(**bad)) == p2r_amd.P2R_ERR_ARG, bad
    ht = torch.zeros(28 * 32, pin_memory=True)
    hh = torch.zeros(9 * 32 * 20, pin_memory=True)
    dt = torch.zeros(28 * 32, device="cuda")
    dh = torch.zeros(9 * 32 * 20, device="cuda")
    s = Shape(1, 32, 20, 32)
    for n in (0, 17):
        assert _status(lambda: ctx.forward_streamed(ht, hh, ht, dt, dh, dt, s, nstreams=n)) == p2r_amd.P2R_ERR_ARG
    assert _status(lambda: ctx.forward_streamed(ht, hh, ht, dt, dh, dt, Shape(1, 33, 20, 32))) == \
        p2r_amd.P2R_ERR_SHAPE
    a = torch.zeros(28 * 64, device="cuda")
    assert _status(lambda: ctx.repack(a, 32, a, 16, 1, 64, 1, 28)) == p2r_amd.P2R_ERR_ARG  # src == dst
    b = torch.zeros(28 * 64, device="cuda")
    assert _status(lambda: ctx.repack(a, 48, b, 16, 1, 64, 1, 28)) == p2r_amd.P2R_ERR_SHAPE  # 48 does not divide 64
    ctx.repack(a, 32, b, 16, 1, 64, 1, 28)  # the context still works


def test_autotuners_refuse_aliased_tracks(ctx):
    """Every autotuner candidate must start from the same input: an output buffer on top of the input is
    refused before anything is timed, and the tuning is left as it was."""
    import p2r_amd
    s = Shape(1, 64, 20, 32)
    t = torch.zeros(28 * 64, device="cuda")
    h = torch.zeros(9 * 64 * 20, device="cuda")
    before = ctx.get_tuning()
    assert _status(lambda: ctx.autotune_forward(t, h, t, s, reps=1)) == p2r_amd.P2R_ERR_ARG
    assert _status(lambda: ctx.autotune_adjoint(t, h, t, s, reps=1)) == p2r_amd.P2R_ERR_ARG
    m = torch.full((64,), 0xfffff, dtype=torch.int32, device="cuda")
    offs = np.array([0, 2], np.int64)
    assert _status(lambda: ctx.autotune_ragged(t, h, t, offs, m, 20, 32, reps=1)) == p2r_amd.P2R_ERR_ARG
    after = ctx.get_tuning()
    assert bytes(before) == bytes(after)


def test_python_mirror_checks_buffers_before_launch(ctx):
    """Buffers shorter than the shape implies, of the wrong dtype, not contiguous, or in host memory are
    refused by the Python mirror before anything reaches the device; the context stays usable."""
    import p2r_amd
    s = Shape(1, 64, 20, 32)
    t = torch.zeros(28 * 64, device="cuda")
    h = torch.zeros(9 * 64 * 20, device="cuda")
    o = torch.zeros(28 * 64, device="cuda")
    with pytest.raises(ValueError):
        ctx.forward(t[:-1], h, o, s)
    with pytest.raises(ValueError):
        ctx.forward(t, h[: 9 * 64 * 19], o, s)
    with pytest.raises(ValueError):
        ctx.forward(t, h, o, s, chi2=torch.zeros(63, device="cuda"))
    with pytest.raises(ValueError):  # per-layer chi2 needs nlayer x n
        ctx.forward(t, h, o, s, chi2=torch.zeros(64, device="cuda"),
                    params=p2r_amd.default_params(options=p2r_amd.P2R_OPT_CHI2_PER_LAYER))
    with pytest.raises(TypeError):
        ctx.forward(t.double(), h, o, s)
    with pytest.raises(TypeError):
        ctx.forward(t.cpu(), h, o, s)
    with pytest.raises(ValueError):
        ctx.forward(torch.zeros(2 * 28 * 64, device="cuda")[::2], h, o, s)
    offs = np.array([0, 2], np.int64)
    m = torch.zeros(64, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError):
        ctx.forward_ragged(t, h, o, offs, m[:63], 20, 32)
    with pytest.raises(TypeError):
        ctx.forward_ragged(t, h, o, offs, m.float(), 20, 32)
    with pytest.raises(ValueError):
        ctx.forward_adjoint(t, h, o, s, grad_hits=torch.zeros(3 * 64 * 19, device="cuda"))
    with pytest.raises(TypeError):
        ctx.forward_adjoint(t, h, o, s, grad_global=torch.zeros(81, device="cuda"))
    ctx.forward(t, h, o, s)  # still usable
    torch.cuda.synchronize()

