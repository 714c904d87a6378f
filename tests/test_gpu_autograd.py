"""GPU tests of the torch.autograd.Function over the adjoint kernel (p2r_amd.autograd, extension
_p2r_torch): gradients equal the C-ABI's adjoint outputs scaled by d loss / d chi2, match the oracle's
FP64 forward-mode gradient, and compose with torch ops for alignment offsets and a field scale."""
import numpy as np
import pytest

from p2r_amd import Shape

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ag():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import p2r_amd.autograd as ag
    return ag


@pytest.fixture(scope="module")
def batch():
    import p2r_amd
    shape = Shape(1, 2048, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    return shape, trk, hits


def _inputs(ag, shape, trk, hits):
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    par, cov, q = ag.split_tracks(dt, shape.bsize)
    hp, hc = ag.split_hits(dh, shape.nlayer, shape.bsize)
    return par.clone().requires_grad_(), cov, q, hp.clone().requires_grad_(), hc


def test_extension_is_the_in_tree_build(ag):
    import os
    assert os.path.dirname(ag._ext.__file__).endswith(os.path.join("p2r-tests_amd", "p2r_amd"))


def test_forward_equals_c_abi_chained(ag, batch):
    """With FP32 gradients, chi2 and fitted state of the autograd path equal p2r_forward (chained, exact)
    bit for bit (the default, mixed, carries the fit in FP64 and returns it rounded to float)."""
    import p2r_amd
    shape, trk, hits = batch
    par, cov, q, hp, hc = _inputs(ag, shape, trk, hits)
    with torch.no_grad():
        chi2, par_out, cov_out = ag.fit(par, cov, q, hp, hc, grad_precision="f32")
    ctx = p2r_amd.Context(0)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    do = torch.empty_like(dt)
    dc = torch.empty(shape.ntrk_total, device="cuda")
    ctx.forward(dt, dh, do, shape, "chained", chi2=dc, precision="exact")
    torch.cuda.synchronize()
    o = do.view(-1, 28, shape.bsize)
    assert torch.equal(chi2.view(torch.int32), dc.view(torch.int32))
    assert torch.equal(par_out.contiguous().view(torch.int32), o[:, :6].contiguous().view(torch.int32))
    assert torch.equal(cov_out.contiguous().view(torch.int32), o[:, 6:27].contiguous().view(torch.int32))


def test_weighted_loss_scales_per_track_gradients(ag, batch):
    """loss = sum_t w_t chi2_t: the gradients are w_t times the unit-weight ones (exactly, per track)."""
    shape, trk, hits = batch
    par, cov, q, hp, hc = _inputs(ag, shape, trk, hits)
    chi2, _, _ = ag.fit(par, cov, q, hp, hc)
    chi2.sum().backward()
    g1p, g1h = par.grad.clone(), hp.grad.clone()
    par.grad, hp.grad = None, None
    w = torch.linspace(0.5, 2.0, shape.ntrk_total, device="cuda")
    chi2, _, _ = ag.fit(par, cov, q, hp, hc)
    (w * chi2).sum().backward()
    wb = w.view(-1, shape.bsize)
    torch.testing.assert_close(par.grad, g1p * wb[:, None, :], rtol=0, atol=0)
    torch.testing.assert_close(hp.grad, g1h * wb[:, None, None, :], rtol=0, atol=0)


def test_f64_gradients_match_dual_oracle(ag, orc, batch):
    shape, trk, hits = batch
    par, cov, q, hp, hc = _inputs(ag, shape, trk, hits)
    kf = ag.default_kfact().cuda().requires_grad_()
    chi2, _, _ = ag.fit(par, cov, q, hp, hc, kfact=kf, grad_precision="f64")
    chi2.sum().backward()
    bs, L = shape.bsize, shape.nlayer
    gp = par.grad.cpu().numpy()
    gh = hp.grad.cpu().numpy()
    T = trk.reshape(-1, 28, bs)
    H = hits.reshape(-1, L, 9, bs)
    gk_total = 0.0
    errs = []
    for t in range(32):
        b, i = divmod(t, bs)
        g, _ = orc.chi2_grad_dual(T[b, :6, i], T[b, 6:27, i], int(T[b, 27, i].view(np.int32)),
                                  np.ascontiguousarray(H[b, :, :, i]), L)
        mine = np.concatenate([gp[b, :, i], gh[b, :, :, i].ravel()]).astype(np.float64)
        errs.append(np.linalg.norm(mine - g[:-1]) / np.linalg.norm(g[:-1]))
    assert max(errs) < 1e-6, max(errs)
    assert kf.grad is not None and torch.isfinite(kf.grad)


def test_alignment_offsets_through_torch_match_grad_global(ag, batch):
    """Per-layer rigid offsets and a field scale composed in torch ops: their gradients at zero equal the
    kernel's FP64 grad_global (dx, dy, dz, dphi per layer; field scale last)."""
    import p2r_amd
    shape, trk, hits = batch
    par, cov, q, hp, hc = _inputs(ag, shape, trk, hits)
    L = shape.nlayer
    offsets = torch.zeros(L, 4, device="cuda", requires_grad=True)
    bscale = torch.ones((), device="cuda", requires_grad=True)
    hp0 = hp.detach()
    chi2, _, _ = ag.fit(par.detach(), cov, q, ag.rigid_transform(hp0, offsets), hc,
                        kfact=ag.default_kfact().cuda() / bscale, grad_precision="f64")
    chi2.double().sum().backward()
    ctx = p2r_amd.Context(0)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    gg = torch.empty(4 * L + 1, dtype=torch.float64, device="cuda")
    ctx.forward_adjoint(dt, dh, torch.empty_like(dt), shape, grad_precision="f64", grad_global=gg)
    torch.cuda.synchronize()
    want = gg.cpu().numpy()
    got = np.concatenate([offsets.grad.cpu().numpy().ravel(), [bscale.grad.item()]])
    # torch reduces the float32 per-hit gradients in float32; the kernel in FP64
    np.testing.assert_allclose(got, want, rtol=2e-4, atol=1e-6 * np.abs(want).max())
