"""GPU tests of p2r_graph_*: a captured forward / uneven-batch / adjoint call replayed several times, with
eager work and synchronisations between the replays (outputs clobbered, other calls on the same context,
including a larger adjoint that regrows the context's workspace), must give exactly the eager call's words
every time."""
import numpy as np
import pytest

from p2r_amd import Shape

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REPLAYS = 3


@pytest.fixture()
def ctx():
    import p2r_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = p2r_amd.Context(0)
    yield c
    c.close()


def _bits(*ts):
    torch.cuda.synchronize()
    return [t.cpu().numpy().view(np.uint32 if t.dtype == torch.float32 else np.uint64).copy() for t in ts]


def _clobber(*ts):
    for t in ts:
        t.fill_(float("nan"))


@pytest.mark.parametrize("sem,per", [("reference", 0), ("chained", 0), ("chained", 7)])
def test_graph_forward_replays_match_eager(ctx, sem, per):
    import p2r_amd
    shape = Shape(3, 1024, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    do, dc = torch.empty_like(dt), torch.empty(shape.ntrk_total, device="cuda")
    ctx.set_tuning(fwd_layers_per_launch=per)
    ctx.forward(dt, dh, do, shape, sem, chi2=dc)
    want = _bits(do, dc)
    g = ctx.graph_forward(dt, dh, do, shape, sem, chi2=dc)
    ctx.set_tuning(fwd_layers_per_launch=0)  # the graph keeps the choices of its capture
    other = Shape(1, 256, 5, 32)
    t2, h2 = (torch.from_numpy(a).cuda() for a in p2r_amd.gen_uniform(other))
    for _ in range(REPLAYS):
        _clobber(do, dc)
        torch.cuda.synchronize()
        g.launch()
        got = _bits(do, dc)
        assert all(np.array_equal(a, b) for a, b in zip(got, want))
        ctx.forward(t2, h2, torch.empty_like(t2), other, "chained", precision="fast")  # eager work between replays
        torch.cuda.synchronize()
    g.close()


@pytest.mark.parametrize("rounds", [-1, 2])
def test_graph_ragged_replays_match_eager(ctx, rounds):
    import p2r_amd
    counts = np.array([700, 33, 1500, 64, 0, 999], np.int32)
    trk, hits, mask, offs = p2r_amd.gen_ragged(counts, 20, 32, kmin=5, kmax=20)
    dt, dh, dm = (torch.from_numpy(a).cuda() for a in (trk, hits, mask.view(np.int32)))
    do = torch.empty_like(dt)
    dc = torch.empty(mask.size, device="cuda")
    ctx.forward_ragged(dt, dh, do, offs, dm, 20, 32, "chained", chi2=dc, tile_rounds=rounds)
    want = _bits(do, dc)
    g = ctx.graph_forward_ragged(dt, dh, do, offs, dm, 20, 32, "chained", chi2=dc, tile_rounds=rounds)
    for _ in range(REPLAYS):
        _clobber(do, dc)
        torch.cuda.synchronize()
        g.launch()
        got = _bits(do, dc)
        assert all(np.array_equal(a, b) for a, b in zip(got, want))
    g.close()


@pytest.mark.parametrize("gp", ["f32", "mixed", "f64"])  # mixed: a third kernel (the FP64 first layers)
def test_graph_adjoint_replays_match_eager(ctx, gp):
    import p2r_amd
    shape = Shape(2, 1024, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    n = shape.ntrk_total
    do = torch.empty_like(dt)
    c2 = torch.empty(n, device="cuda")
    gpar = torch.empty(6 * n, device="cuda")
    ghit = torch.empty(3 * n * shape.nlayer, device="cuda")
    gk = torch.empty(n, device="cuda")
    gg = torch.empty(4 * shape.nlayer + 1, device="cuda", dtype=torch.float64)
    outs = (do, c2, gpar, ghit, gk, gg)
    kw = dict(grad_precision=gp, chi2=c2, grad_par=gpar, grad_hits=ghit, grad_kfact=gk, grad_global=gg)
    ctx.forward_adjoint(dt, dh, do, shape, **kw)
    want = _bits(*outs)
    g = ctx.graph_forward_adjoint(dt, dh, do, shape, **kw)
    # A larger eager adjoint regrows the context's workspace; the graph has its own.
    big = Shape(4, 2048, 20, 32)
    tb, hb = (torch.from_numpy(a).cuda() for a in p2r_amd.gen_uniform(big))
    nb = big.ntrk_total
    for _ in range(REPLAYS):
        _clobber(*outs)
        torch.cuda.synchronize()
        g.launch()
        got = _bits(*outs)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        ctx.forward_adjoint(tb, hb, torch.empty_like(tb), big, grad_precision="f64",
                            grad_global=torch.empty(4 * big.nlayer + 1, device="cuda", dtype=torch.float64),
                            grad_kfact=torch.empty(nb, device="cuda"))
        torch.cuda.synchronize()
    g.close()


def test_graph_of_empty_batch_and_bad_arguments(ctx):
    import p2r_amd
    shape = Shape(0, 1024, 20, 32)
    dt = torch.empty(32, device="cuda")
    g = ctx.graph_forward(dt, dt, dt, shape, "reference")
    g.launch()  # nothing captured, nothing launched
    g.close()
    with pytest.raises(p2r_amd.P2RError):
        ctx.graph_forward(dt, dt, dt, Shape(1, 100, 20, 32), "reference")  # ntrks not a multiple of bsize
    with pytest.raises(p2r_amd.P2RError):
        ctx.graph_forward_adjoint(dt, dt, dt, Shape(1, 64, 20, 32),
                                  params=p2r_amd.default_params("exact", options=1))
