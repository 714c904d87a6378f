"""GPU tests of the hand-written reverse-mode adjoint (p2r_forward_adjoint).

Reference gradients: the oracle's forward-mode Dual<16> FP64 differentiation
of the same algorithm, and FP64 central finite differences (h = 1e-6 relative)
of the oracle's FP64 total chi2.  Error measure per track: the norm-wise
relative error ||g - g_ref|| / ||g_ref|| over the 6 + 3*nlayer + 1 gradient
components (initial parameters, hit coordinates, kfact).
"""
import numpy as np
import pytest

import oracle
from p2r_amd import Shape

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    import p2r_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return p2r_amd.Context(0)


def run_adjoint(ctx, trk, hits, shape, gp="f32", precision="exact", cs=0, with_global=False):
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    do = torch.empty_like(dt)
    n = shape.ntrk_total
    c2 = torch.empty(n, device="cuda")
    gpar = torch.empty(shape.nblk * 6 * shape.bsize, device="cuda")
    ghit = torch.empty(shape.nblk * shape.nlayer * 3 * shape.bsize, device="cuda")
    gk = torch.empty(n, device="cuda")
    gg = torch.empty(4 * shape.nlayer + 1, dtype=torch.float64, device="cuda") if with_global else None
    ctx.forward_adjoint(dt, dh, do, shape, grad_precision=gp, precision=precision, checkpoint_stride=cs, chi2=c2,
                        grad_par=gpar, grad_hits=ghit, grad_kfact=gk, grad_global=gg)
    torch.cuda.synchronize()
    bs, L = shape.bsize, shape.nlayer
    gp_ = gpar.cpu().numpy().reshape(-1, 6, bs).transpose(0, 2, 1).reshape(-1, 6)
    gh_ = ghit.cpu().numpy().reshape(-1, L, 3, bs).transpose(0, 3, 1, 2).reshape(-1, 3 * L)
    grads = np.hstack([gp_, gh_, gk.cpu().numpy()[:, None]]).astype(np.float64)
    return (do.cpu().numpy(), c2.cpu().numpy(), grads, gg.cpu().numpy() if with_global else None)


def _track(trk, hits, shape, t):
    bs, L = shape.bsize, shape.nlayer
    T = trk.reshape(-1, 28, bs)
    H = hits.reshape(-1, L, 9, bs)
    b, i = divmod(t, bs)
    return (T[b, :6, i].copy(), T[b, 6:27, i].copy(), int(T[b, 27, i].view(np.int32)),
            np.ascontiguousarray(H[b, :, :, i]))


def _fd_grad(orc, par, cov, q, h, L):
    x = np.concatenate([par.astype(np.float64), h[:, :3].astype(np.float64).ravel()])
    kf = float(orc.kfact_tbb())

    def f(xx, kk=kf):
        return orc.chi2_f64(xx[:6], cov, q, xx[6:], h, L, kk)
    g = np.zeros(len(x) + 1)
    for j in range(len(x)):
        d = 1e-6 * max(abs(x[j]), 1e-3)
        xp, xm = x.copy(), x.copy()
        xp[j] += d
        xm[j] -= d
        g[j] = (f(xp) - f(xm)) / (2 * d)
    d = 1e-6 * abs(kf)
    g[-1] = (f(x, kf + d) - f(x, kf - d)) / (2 * d)
    return g


@pytest.fixture(scope="module")
def batch():
    import p2r_amd
    shape = Shape(2, 1024, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    return shape, trk, hits


def _errors(orc, grads, batch, ref="dual", ntracks=64):
    shape, trk, hits = batch
    errs = []
    for t in range(ntracks):
        par, cov, q, h = _track(trk, hits, shape, t)
        if ref == "dual":
            gref, _ = orc.chi2_grad_dual(par, cov, q, h, shape.nlayer)
        else:
            gref = _fd_grad(orc, par, cov, q, h, shape.nlayer)
        errs.append(np.linalg.norm(grads[t] - gref) / np.linalg.norm(gref))
    return np.array(errs)


def test_adjoint_forward_outputs_equal_p2r_forward(ctx, batch):
    """The adjoint's forward pass (exact, F32) reproduces p2r_forward chained bit for bit, chi2 included."""
    shape, trk, hits = batch
    out, c2, _, _ = run_adjoint(ctx, trk, hits, shape, "f32", "exact")
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    do = torch.empty_like(dt)
    dc = torch.empty(shape.ntrk_total, device="cuda")
    ctx.forward(dt, dh, do, shape, "chained", chi2=dc, precision="exact")
    torch.cuda.synchronize()
    assert np.array_equal(out.view(np.uint32), do.cpu().numpy().view(np.uint32))
    assert np.array_equal(c2.view(np.uint32), dc.cpu().numpy().view(np.uint32))


def test_adjoint_fast_forward_outputs_meet_the_fast_criterion(orc, ctx, batch):
    """With fast arithmetic the adjoint's forward pass is its own compilation of the fused formulas (the
    compiler fuses multiply-adds per kernel), so its outputs are held to the fast-mode criterion of the
    plain forward (tests/parity.py::compare_statistical: as close to the IEEE-strict oracle as the
    reference's own fast-math build) rather than to p2r_forward's bits."""
    import parity
    shape, trk, hits = batch
    out, _, _, _ = run_adjoint(ctx, trk, hits, shape, "f32", "fast")
    cpu = orc.run_f32(trk, hits, shape.nblk, shape.nlayer, shape.bsize, semantics=1)
    alt = oracle.Oracle("liboracle_bench.so").run_f32(trk, hits, shape.nblk, shape.nlayer, shape.bsize, semantics=1)
    par64, cov64, _ = orc.run_f64(trk, hits, shape.nblk, shape.nlayer, shape.bsize, semantics=1)
    res = parity.compare_statistical(out.ravel(), cpu, alt, shape.bsize, par64, cov64)
    print("fast adjoint forward pass", res.get("frac_within_rtol_of_cpu"), res.get("ok"), flush=True)
    assert res["ok"], res


def test_adjoint_f64_matches_dual_and_fd(orc, ctx, batch):
    _, _, grads, _ = run_adjoint(ctx, *batch[1:], batch[0], "f64")
    e_dual = _errors(orc, grads, batch, "dual")
    e_fd = _errors(orc, grads, batch, "fd", ntracks=16)
    print(f"f64 adjoint: vs Dual max {e_dual.max():.2e}; vs FD max {e_fd.max():.2e}", flush=True)
    assert e_dual.max() < 1e-6
    assert e_fd.max() < 1e-4


@pytest.mark.parametrize("gp,precision,median_bound,max_bound", [("mixed", "exact", 1e-3, 1e-3),
                                                                  ("f32", "exact", 1.5e-3, 1.5e-2),
                                                                  ("f32", "fast", 1.5e-3, 1.5e-2)])
def test_adjoint_fp32_gradients_vs_fp64(orc, ctx, batch, gp, precision, median_bound, max_bound):
    """FP32 arithmetic against FP64 FD on 64 tracks.  Carrying the state in FP64, with the first two
    layers' reverse in FP64 (mixed, the default gradient precision), keeps every track within 1e-3; the
    chained fit's all-FP32 states carry the Kalman update's cancellation, so the all-FP32 gradient is
    bounded per track only to ~1e-2, in either arithmetic of the forward pass."""
    _, _, grads, _ = run_adjoint(ctx, *batch[1:], batch[0], gp, precision)
    e = _errors(orc, grads, batch, "fd")
    print(f"{gp} ({precision}) adjoint vs FP64 FD on 64 tracks: median {np.median(e):.2e} "
          f"p90 {np.quantile(e, 0.9):.2e} max {e.max():.2e}", flush=True)
    assert np.median(e) < median_bound and e.max() < max_bound


def test_checkpoint_stride_does_not_change_gradients(ctx, batch):
    """Exact FP32: the segments' recomputed inputs are the forward pass's states bit for bit (no contraction,
    so both kernels round every operation alike), and the gradients are the same words at every stride."""
    shape, trk, hits = batch
    ref = run_adjoint(ctx, trk, hits, shape, "f32", cs=4)[2]
    for cs in (1, 2, 8):
        g = run_adjoint(ctx, trk, hits, shape, "f32", cs=cs)[2]
        assert np.array_equal(g.view(np.uint64), ref.view(np.uint64)), cs


@pytest.mark.parametrize("gp,precision,tol", [("f32", "fast", 1.5e-2), ("mixed", "exact", 1e-5), ("f64", "exact", 1e-8)])
def test_checkpoint_stride_with_fused_arithmetic(ctx, batch, gp, precision, tol):
    """Where the recompute fuses multiply-adds (fast FP32, and the FP64 states of MIXED and F64), the
    compiler fuses them per kernel, so the forward pass and a segment's recompute may round a state
    differently in its last bit; the gradients then agree across strides to the arithmetic's precision
    (normwise per track), not word for word.  The chained fit's ill-conditioned updates amplify a last-bit
    difference of an FP32 state to ~2e-3 of a gradient (measured), the scale of the FP32 gradient's own
    error against FP64 (bounded at 1.5e-2 above); FP64 states differ by ~1e-12 to 2e-9."""
    shape, trk, hits = batch
    ref = run_adjoint(ctx, trk, hits, shape, gp, precision, cs=4)[2]
    for cs in (1, 2, 8):
        g = run_adjoint(ctx, trk, hits, shape, gp, precision, cs=cs)[2]
        e = np.linalg.norm(g - ref, axis=1) / np.linalg.norm(ref, axis=1)
        print(f"{gp} ({precision}) stride {cs} vs 4: max per-track normwise difference {e.max():.2e}", flush=True)
        assert e.max() < tol, cs


@pytest.mark.parametrize("gp", ["f64", "mixed"])
def test_global_gradients_reduce_the_per_hit_gradients(ctx, batch, gp):
    """grad_global = sums over tracks of the per-hit gradients (dx, dy, dz, dphi per layer) and of
    -kfact * d/dkfact (field scale), reduced in FP64 in a fixed order: bit-reproducible run to run.  With
    mixed, the FP64 reverse of the first two layers writes their entries and the field scale's."""
    shape, trk, hits = batch
    _, _, grads, gg = run_adjoint(ctx, trk, hits, shape, gp, with_global=True)
    _, _, _, gg2 = run_adjoint(ctx, trk, hits, shape, gp, with_global=True)
    assert np.array_equal(gg.view(np.uint64), gg2.view(np.uint64))
    L, bs = shape.nlayer, shape.bsize
    H = hits.reshape(-1, L, 9, bs).transpose(0, 3, 1, 2).reshape(-1, L, 9).astype(np.float64)
    gh = grads[:, 6:6 + 3 * L].reshape(-1, L, 3)
    want = np.zeros(4 * L + 1)
    want[0:4 * L:4] = gh[:, :, 0].sum(axis=0)
    want[1:4 * L:4] = gh[:, :, 1].sum(axis=0)
    want[2:4 * L:4] = gh[:, :, 2].sum(axis=0)
    want[3:4 * L:4] = (-H[:, :, 1] * gh[:, :, 0] + H[:, :, 0] * gh[:, :, 1]).sum(axis=0)
    import p2r_amd
    want[4 * L] = -p2r_amd.default_params().kfact * grads[:, -1].sum()
    np.testing.assert_allclose(gg, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())


def test_alignment_gradient_matches_finite_difference_of_total_chi2(orc, ctx):
    """d(sum chi2)/d(dx of layer 7) and d/d(dphi of layer 12) and d/d(b_scale) on 64 tracks against FP64
    central differences of the oracle's total chi2 with the hits moved / the field scaled."""
    import p2r_amd
    shape = Shape(1, 64, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    _, _, _, gg = run_adjoint(ctx, trk, hits, shape, "f64", with_global=True)
    kf = float(orc.kfact_tbb())

    def total(dx7=0.0, dphi12=0.0, bscale=1.0):
        s = 0.0
        for t in range(64):
            par, cov, q, h = _track(trk, hits, shape, t)
            hp = h[:, :3].astype(np.float64).copy()
            hp[7, 0] += dx7
            c, sn = np.cos(dphi12), np.sin(dphi12)
            x, y = hp[12, 0], hp[12, 1]
            hp[12, 0], hp[12, 1] = c * x - sn * y, sn * x + c * y
            s += orc.chi2_f64(par.astype(np.float64), cov, q, hp.ravel(), h, 20, kf / bscale)
        return s
    for idx, kw, d in ((4 * 7 + 0, "dx7", 1e-5), (4 * 12 + 3, "dphi12", 1e-7), (4 * 20, "bscale", 1e-7)):
        fd = (total(**{kw: d}) - total(**{kw: -d})) / (2 * d) if kw != "bscale" else \
             (total(bscale=1 + d) - total(bscale=1 - d)) / (2 * d)
        print(f"global[{idx}] ({kw}): adjoint {gg[idx]:.6e} FD {fd:.6e}", flush=True)
        assert abs(gg[idx] - fd) <= 1e-4 * abs(fd) + 1e-6, (kw, gg[idx], fd)


def test_global_gradients_skip_diverged_tracks(ctx, batch):
    """A track whose fit is not finite (here: NaN initial 1/pt) is left out of grad_global; the other
    tracks' sums are unchanged (compared with the same batch where that track's hits are zeroed out of the
    expected sum)."""
    shape, trk, hits = batch
    bad = 37
    t2 = trk.copy()
    T = t2.reshape(-1, 28, shape.bsize)
    T[bad // shape.bsize, 3, bad % shape.bsize] = np.nan
    _, c2, grads, gg = run_adjoint(ctx, t2, hits, shape, "f64", with_global=True)
    assert not np.isfinite(c2[bad]) and np.isfinite(np.delete(c2, bad)).all()
    assert np.isfinite(gg).all()
    _, _, grads0, gg0 = run_adjoint(ctx, trk, hits, shape, "f64", with_global=True)
    L, bs = shape.nlayer, shape.bsize
    gh = grads0[bad, 6:6 + 3 * L].reshape(L, 3)
    H = hits.reshape(-1, L, 9, bs)[bad // bs, :, :, bad % bs].astype(np.float64)
    contrib = np.zeros(4 * L + 1)
    contrib[0:4 * L:4], contrib[1:4 * L:4], contrib[2:4 * L:4] = gh[:, 0], gh[:, 1], gh[:, 2]
    contrib[3:4 * L:4] = -H[:, 1] * gh[:, 0] + H[:, 0] * gh[:, 1]
    import p2r_amd
    contrib[4 * L] = -p2r_amd.default_params().kfact * grads0[bad, -1]
    np.testing.assert_allclose(gg, gg0 - contrib, rtol=1e-6, atol=1e-9 * np.abs(gg0).max())
