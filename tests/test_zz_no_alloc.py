"""p2r_forward_adjoint_noalloc: the fit + adjoint with its scratch in the library's static device array.

Called through the C-ABI (ctypes), its outputs are compared word for word with p2r_forward_adjoint's on
the same inputs -- both run the same kernels, so the results are identical -- on two calls in a row, the
second with a smaller batch (the array keeps the first call's contents), then a request larger than the
array is refused with P2R_ERR_SHAPE before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

from p2r_amd import Shape

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    import p2r_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return p2r_amd.Context(0)


def _buffers(shape, with_global):
    n = shape.ntrk_total
    return dict(out=torch.full((shape.trk_floats,), float("nan"), device="cuda"),
                chi2=torch.full((n,), float("nan"), device="cuda"),
                gpar=torch.full((shape.nblk * 6 * shape.bsize,), float("nan"), device="cuda"),
                ghit=torch.full((shape.nblk * shape.nlayer * 3 * shape.bsize,), float("nan"), device="cuda"),
                gk=torch.full((n,), float("nan"), device="cuda"),
                gg=torch.full((4 * shape.nlayer + 1,), float("nan"), dtype=torch.float64, device="cuda")
                if with_global else None)


def _call(ctx, fn, dt, dh, shape, gp, cs, with_global):
    import p2r_amd
    from p2r_amd import _capi
    b = _buffers(shape, with_global)
    p = p2r_amd.default_params()
    sh = _capi.p2r_shape(*shape.tuple())
    P = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    st = getattr(_capi.lib(), fn)(ctx._h, P(dt), P(dh), P(b["out"]), C.byref(sh), C.byref(p),
                                  p2r_amd.GRAD_PRECISION[gp], cs, P(b["chi2"]), P(b["gpar"]), P(b["ghit"]),
                                  P(b["gk"]), P(b["gg"]), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, {k: (v.cpu().numpy() if v is not None else None) for k, v in b.items()}


def _same(a, b):
    ua = a.view(np.uint64 if a.dtype == np.float64 else np.uint32)
    ub = b.view(np.uint64 if b.dtype == np.float64 else np.uint32)
    return bool(np.array_equal(ua, ub))


@pytest.mark.parametrize("gp,cs", [("f32", 4), ("mixed", 2), ("f64", 4)])  # f64 4: segment states in scratch
def test_noalloc_equals_adjoint_on_two_calls(ctx, gp, cs):
    import p2r_amd
    for shape in (Shape(3, 8192, 20, 32), Shape(1, 2048, 20, 32)):  # the second smaller than the first
        print(f"noalloc {gp} stride {cs} {shape}", flush=True)
        trk, hits = p2r_amd.gen_uniform(shape)
        dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
        st0, want = _call(ctx, "p2r_forward_adjoint", dt, dh, shape, gp, cs, True)
        st1, got = _call(ctx, "p2r_forward_adjoint_noalloc", dt, dh, shape, gp, cs, True)
        assert st0 == 0 and st1 == 0, (st0, st1, p2r_amd._capi.lib().p2r_last_error(ctx._h))
        for k in want:
            assert np.isfinite(want[k]).any(), k
            assert _same(got[k], want[k]), (shape, k)
        # without the global sums (no partials in the scratch)
        st2, got2 = _call(ctx, "p2r_forward_adjoint_noalloc", dt, dh, shape, gp, cs, False)
        assert st2 == 0
        for k in ("out", "chi2", "gpar", "ghit", "gk"):
            assert _same(got2[k], want[k]), (shape, k)


def test_noalloc_refuses_more_than_its_scratch(ctx):
    import p2r_amd
    cap = p2r_amd.P2R_NOALLOC_SCRATCH_BYTES
    # stride 1 in FP64: 19 checkpoints x 27 states x 8 bytes per track, far beyond the array at 32768 tracks
    shape = Shape(1, 32768, 20, 32)
    assert 19 * 27 * 8 * shape.ntrk_total > cap
    dt = torch.zeros(shape.trk_floats, device="cuda")
    dh = torch.zeros(shape.hit_floats, device="cuda")
    st, got = _call(ctx, "p2r_forward_adjoint_noalloc", dt, dh, shape, "f64", 1, True)
    assert st == p2r_amd.P2R_ERR_SHAPE
    assert b"static" in p2r_amd._capi.lib().p2r_last_error(ctx._h)
    assert np.isnan(got["out"]).all() and np.isnan(got["gg"]).all()  # nothing was launched
    # the context and the entry point still work
    small = Shape(1, 256, 20, 32)
    trk, hits = p2r_amd.gen_uniform(small)
    st, _ = _call(ctx, "p2r_forward_adjoint_noalloc", torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda(),
                  small, "f32", 4, True)
    assert st == 0
