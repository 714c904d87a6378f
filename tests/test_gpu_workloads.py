"""GPU tests at the benchmark workloads' full sizes, each against the oracle:

  * uneven batch: 500 events (lognormal track counts, median 9600, sigma 0.5, clipped to 500..40000;
    each track with k ~ U{5..20} present layers of 20), chained, the pair-stream schedule captured in a
    HIP graph, at Matriplex widths 16, 32 and 64: every word of the output and the per-track chi2 equal to
    the oracle's masked run;
  * differentiable fit: 100 x 9600 x 20, checkpoints every 4 layers: the fitted tracks and chi2 of the
    FP32 pass equal the oracle's chained run word for word; per-track gradients of 64 tracks drawn at
    random (seeded) from the whole batch against the oracle's FP64 forward-mode Dual<16> gradient and
    FP64 central differences (h = 1e-6 relative);
  * alignment gradient: 1 000 000 tracks x 40 layers with hidden per-layer offsets and field scale: FP64
    per-track gradients of 64 random tracks against Dual<16>, the global sums against the FP64 sums of
    the per-hit gradients, and both bit-reproducible from run to run.

Error measure of a gradient, per track: ||g - g_ref|| / ||g_ref|| over its 6 + 3 * nlayer + 1 components
(initial parameters, hit coordinates, kfact).  Each case prints its name and sizes before it runs.
"""
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


def _say(msg):
    print(msg, flush=True)


def _same_words(a, b):
    a = np.asarray(a, np.float32).ravel()
    b = np.asarray(b, np.float32).ravel()
    return int(((a.view(np.uint32) != b.view(np.uint32)) & ~(np.isnan(a) & np.isnan(b))).sum())


# ---- uneven batch, 500 events -----------------------------------------------------------------------------

@pytest.fixture(scope="module")
def counts_500():
    import p2r_amd
    return p2r_amd.lognormal_counts(500, median=9600, sigma=0.5, lo=500, hi=40000)


@pytest.mark.parametrize("bsize", [16, 32, 64])
def test_uneven_500_events_graph_pair_stream_matches_oracle(orc, ctx, counts_500, bsize):
    import p2r_amd
    trk, hits, mask, offs = p2r_amd.gen_ragged(counts_500, 20, bsize, kmin=5, kmax=20)
    nblk = int(offs[-1])
    updates = int(np.unpackbits(mask.view(np.uint8)).sum())
    _say(f"uneven 500 events bsize {bsize}: {int(counts_500.sum())} tracks, {mask.size} slots, {updates} updates")
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    dm = torch.from_numpy(mask.view(np.int32)).cuda()
    do = torch.zeros_like(dt)
    dc = torch.full((mask.size,), float("nan"), device="cuda")
    g = ctx.graph_forward_ragged(dt, dh, do, offs, dm, 20, bsize, "chained", chi2=dc, precision="exact",
                                 tile_rounds=-2)
    cpu, c2c = orc.run_f32(trk, hits, nblk, 20, bsize, semantics=1, chi2=True, mask=mask, cr_trig=True)
    del hits
    for replay in range(2):
        do.zero_()
        dc.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        n_state = _same_words(do.cpu().numpy(), cpu)
        n_chi2 = _same_words(dc.cpu().numpy(), c2c)
        _say(f"  replay {replay}: {n_state} output words and {n_chi2} chi2 words differ from the oracle")
        assert n_state == 0 and n_chi2 == 0
    g.close()


# ---- differentiable fit, 100 x 9600 x 20, checkpoint stride 4 ------------------------------------------------

def _adjoint(ctx, dt, dh, shape, gp, cs=4, with_global=False, precision="exact"):
    n = shape.ntrk_total
    do = torch.empty_like(dt)
    c2 = torch.empty(n, device="cuda")
    gpar = torch.empty(shape.nblk * 6 * shape.bsize, device="cuda")
    ghit = torch.empty(shape.nblk * shape.nlayer * 3 * shape.bsize, device="cuda")
    gk = torch.empty(n, device="cuda")
    gg = torch.empty(4 * shape.nlayer + 1, dtype=torch.float64, device="cuda") if with_global else None
    ctx.forward_adjoint(dt, dh, do, shape, grad_precision=gp, precision=precision, checkpoint_stride=cs, chi2=c2,
                        grad_par=gpar, grad_hits=ghit, grad_kfact=gk, grad_global=gg)
    torch.cuda.synchronize()
    return do, c2, gpar, ghit, gk, gg


def _track_grads(shape, gpar, ghit, gk, tracks):
    """Gradient rows [d/dpar (6), d/dhit (3 per layer), d/dkfact] of the given flat track indices."""
    bs, L = shape.bsize, shape.nlayer
    P = gpar.view(-1, 6, bs)
    H = ghit.view(-1, L, 3, bs)
    t = torch.as_tensor(tracks, device=P.device)
    b, i = t // bs, t % bs
    gp_ = P[b, :, i]
    gh_ = H[b, :, :, i].reshape(len(tracks), 3 * L)
    return torch.cat([gp_, gh_, gk[t][:, None]], 1).double().cpu().numpy()


def _track_inputs(trk, hits, shape, t):
    bs, L = shape.bsize, shape.nlayer
    T = trk.reshape(-1, 28, bs)
    H = hits.reshape(-1, L, 9, bs)
    b, i = divmod(int(t), bs)
    return (T[b, :6, i].copy(), T[b, 6:27, i].copy(), int(T[b, 27, i].view(np.int32)),
            np.ascontiguousarray(H[b, :, :, i]))


def _fd_grad(orc, par, cov, q, h, L):
    """FP64 central differences (h = 1e-6 relative) of the oracle's FP64 total chi2."""
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
def fit_100():
    import p2r_amd
    shape = Shape(100, 9600, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    rng = np.random.default_rng(0x4D495844)
    tracks = np.sort(rng.choice(shape.ntrk_total, 64, replace=False))
    return shape, trk, hits, tracks


@pytest.fixture(scope="module")
def fit_100_refs(orc, fit_100):
    """Dual<16> and central-difference gradients of the 64 sampled tracks (finite fits only)."""
    shape, trk, hits, tracks = fit_100
    dual, fd = [], []
    for t in tracks:
        par, cov, q, h = _track_inputs(trk, hits, shape, t)
        dual.append(orc.chi2_grad_dual(par, cov, q, h, shape.nlayer)[0])
        fd.append(_fd_grad(orc, par, cov, q, h, shape.nlayer))
    return np.array(dual), np.array(fd)


def test_fit_100x9600x20_stride4_outputs_match_oracle(orc, ctx, fit_100):
    shape, trk, hits, _ = fit_100
    _say(f"fit + adjoint {shape}, stride 4, exact FP32: outputs against the oracle's chained run")
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    do, c2, *_ = _adjoint(ctx, dt, dh, shape, "f32")
    cpu, c2c = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=1, chi2=True, cr_trig=True)
    n_state, n_chi2 = _same_words(do.cpu().numpy(), cpu), _same_words(c2.cpu().numpy(), c2c)
    _say(f"  {n_state} output words and {n_chi2} chi2 words differ")
    assert n_state == 0 and n_chi2 == 0


@pytest.mark.parametrize("gp,max_bound,median_bound", [("mixed", 1e-3, 5e-4), ("f64", 1e-5, 1e-6),
                                                       ("f32", 1.5e-2, 1.5e-3)])
def test_fit_100x9600x20_stride4_gradients(ctx, fit_100, fit_100_refs, gp, max_bound, median_bound):
    """FP64 within 1e-5 of FP64 central differences (their own truncation) on every sampled track.  The
    default gradient precision (mixed: FP64 state, FP32 reverse but for the first two layers') within 1e-3
    on every sampled track, the gradient accuracy the adjoint is built to; all-FP32 carries the FP32 update's cancellation in its
    fitted states (a quarter of the tracks above 1e-3)."""
    shape, trk, hits, tracks = fit_100
    dual, fd = fit_100_refs
    _say(f"fit + adjoint {shape}, stride 4, {gp}: 64 random tracks {tracks[:4]}... against Dual and FD")
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    _, _, gpar, ghit, gk, _ = _adjoint(ctx, dt, dh, shape, gp)
    g = _track_grads(shape, gpar, ghit, gk, tracks)
    e_fd = np.linalg.norm(g - fd, axis=1) / np.linalg.norm(fd, axis=1)
    e_dual = np.linalg.norm(g - dual, axis=1) / np.linalg.norm(dual, axis=1)
    _say(f"  vs FD: median {np.median(e_fd):.2e} p90 {np.quantile(e_fd, 0.9):.2e} max {e_fd.max():.2e}; "
         f"vs Dual: median {np.median(e_dual):.2e} max {e_dual.max():.2e}")
    assert np.isfinite(g).all()
    assert e_fd.max() < max_bound and np.median(e_fd) < median_bound
    if gp == "f64":
        assert e_dual.max() < 1e-6


# ---- alignment gradient, 1e6 tracks x 40 layers -------------------------------------------------------------

@pytest.fixture(scope="module")
def align_1m():
    import p2r_amd
    nl = 40
    shape = Shape(1, 1000000, nl, 32)
    rng = np.random.default_rng(20240905)
    offsets = np.zeros((nl, 4), np.float32)
    offsets[:, :3] = rng.normal(0, 50e-4, (nl, 3))  # 50 um in cm
    offsets[:, 3] = rng.normal(0, 1e-4, nl)
    bscale = float(1 + rng.normal(0, 1e-3))
    cfg = p2r_amd.gen_config(r0=20.0, dr=2.5, layer_offsets=offsets.ravel(), b_scale_truth=bscale)
    trk, hits = p2r_amd.gen_uniform(shape, cfg)
    tracks = np.sort(np.random.default_rng(0x414C49).choice(shape.ntrk_total, 64, replace=False))
    return shape, trk, hits, tracks


def test_alignment_1m_x40_gradients(orc, ctx, align_1m):
    import p2r_amd
    shape, trk, hits, tracks = align_1m
    L, bs = shape.nlayer, shape.bsize
    _say(f"alignment {shape}: FP64 adjoint, 64 random tracks against Dual, global sums, reproducibility")
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    _, c2, gpar, ghit, gk, gg = _adjoint(ctx, dt, dh, shape, "f64", with_global=True)
    gg = gg.cpu().numpy()
    # per-track gradients of 64 random tracks against the oracle's forward-mode FP64 gradient
    g = _track_grads(shape, gpar, ghit, gk, tracks)
    dual = np.array([orc.chi2_grad_dual(*_track_inputs(trk, hits, shape, t), L)[0] for t in tracks])
    e = np.linalg.norm(g - dual, axis=1) / np.linalg.norm(dual, axis=1)
    _say(f"  per-track vs Dual<16>: median {np.median(e):.2e} max {e.max():.2e}")
    assert e.max() < 1e-6
    # global sums = FP64 sums over the tracks with a finite total chi2 of the per-hit gradients (dx, dy, dz,
    # dphi = -y d/dx + x d/dy per layer) and of -kfact d/dkfact (field scale).  The per-hit gradients come
    # back in float: each term is off by at most 2^-24 of itself, so the sums agree to 1e-6 of sum |term|.
    ok = torch.isfinite(c2)
    H = dh.view(-1, L, 9, bs)
    Gh = ghit.view(-1, L, 3, bs)
    want = np.zeros(4 * L + 1)
    scale = np.zeros(4 * L + 1)
    okb = ok.view(-1, bs)
    for l in range(L):
        gx, gy, gz = (Gh[:, l, c, :].double() for c in range(3))
        hx, hy = H[:, l, 0, :].double(), H[:, l, 1, :].double()
        terms = (gx, gy, gz, -hy * gx + hx * gy)
        for c, t in enumerate(terms):
            t = torch.where(okb, t, torch.zeros_like(t))
            want[4 * l + c] = float(t.sum())
            scale[4 * l + c] = float(t.abs().sum())
    kf = p2r_amd.default_params().kfact
    tk = torch.where(ok, -kf * gk.double(), torch.zeros_like(gk, dtype=torch.float64))
    want[4 * L], scale[4 * L] = float(tk.sum()), float(tk.abs().sum())
    dev = np.abs(gg - want)
    _say(f"  global sums: max |gg - sum| / sum|term| = {(dev / scale).max():.2e}; "
         f"{int((~ok).sum())} diverged tracks left out")
    assert np.all(dev <= 1e-6 * scale + 1e-12)
    # bit-reproducible: the same call twice, and the FP32 gradient precision twice
    gg2 = _adjoint(ctx, dt, dh, shape, "f64", with_global=True)[5].cpu().numpy()
    assert np.array_equal(gg.view(np.uint64), gg2.view(np.uint64))
    a = _adjoint(ctx, dt, dh, shape, "f32", with_global=True)[5].cpu().numpy()
    b = _adjoint(ctx, dt, dh, shape, "f32", with_global=True)[5].cpu().numpy()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
