"""GPU tests of the launch choices (p2r_set_tuning) and the autotuners: every choice gives bit-identical
results (the per-lane arithmetic does not depend on workgroup size, layer chunking or Matriplex width),
the autotuners pick one of their candidates and leave the context set to it, and the width repack is a
permutation that round-trips."""
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


def _fwd(ctx, dt, dh, shape, sem, precision="exact"):
    do = torch.empty_like(dt)
    dc = torch.empty(shape.ntrk_total, device="cuda")
    ctx.forward(dt, dh, do, shape, sem, chi2=dc, precision=precision)
    torch.cuda.synchronize()
    return do.cpu().numpy().view(np.uint32), dc.cpu().numpy().view(np.uint32)


@pytest.fixture(scope="module")
def batch():
    import p2r_amd
    shape = Shape(3, 1024, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    return shape, torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()


def _oracle_words(orc, dt, dh, shape, sem):
    """The oracle's IEEE-strict FP32 run (correctly rounded sin/cos) on the same inputs, as words."""
    o, c = orc.run_f32(dt.cpu().numpy(), dh.cpu().numpy(), shape.nblk, shape.nlayer, shape.bsize,
                       semantics=0 if sem == "reference" else 1, chi2=True, cr_trig=True)
    return o.view(np.uint32), c.view(np.uint32)


def _equal_words(a, b):
    """Word-for-word equality, NaN equal to NaN whatever its payload."""
    fa, fb = a.view(np.float32), b.view(np.float32)
    return bool(np.all((a == b) | (np.isnan(fa) & np.isnan(fb))))


@pytest.mark.parametrize("sem", ["reference", "chained"])
@pytest.mark.parametrize("per", [0, 7, 1])
def test_in_place_forward_matches_oracle(ctx, orc, batch, sem, per):
    """trk_out == trk_in, with the layers in one launch or split over launches of 7 and of 1 layer: the
    input must survive until the last launch has read it (in reference semantics every launch restarts
    from it).  Bit-identical to the oracle, state, charge and chi2."""
    shape, dt, dh = batch
    want_o, want_c = _oracle_words(orc, dt, dh, shape, sem)
    ctx.set_tuning(fwd_layers_per_launch=per)
    x = dt.clone()
    dc = torch.full((shape.ntrk_total,), float("nan"), device="cuda")
    ctx.forward(x, dh, x, shape, sem, chi2=dc, precision="exact")
    torch.cuda.synchronize()
    got_o, got_c = x.cpu().numpy().view(np.uint32), dc.cpu().numpy().view(np.uint32)
    print(f"in place {sem} layers/launch={per}: {int((got_o != want_o).sum())} words differ", flush=True)
    assert _equal_words(got_o, want_o) and _equal_words(got_c, want_c)


def test_tuning_roundtrip_and_validation(ctx):
    import p2r_amd
    t = ctx.get_tuning()
    assert t.as_dict() == {"fwd_threads": 256, "fwd_layers_per_launch": 0, "ragged_schedule": -2,
                           "ragged_threads": 256, "checkpoint_stride": 4, "fwd_waves": 0}
    ctx.set_tuning(fwd_threads=128, checkpoint_stride=2)
    assert ctx.get_tuning().fwd_threads == 128 and ctx.get_tuning().checkpoint_stride == 2
    for bad in ({"fwd_threads": 96}, {"ragged_schedule": 3}, {"checkpoint_stride": 3}, {"fwd_layers_per_launch": -1},
                {"fwd_waves": 2}):
        with pytest.raises(p2r_amd.P2RError):
            ctx.set_tuning(**bad)
    assert ctx.get_tuning().fwd_threads == 128  # a rejected setting changes nothing


@pytest.mark.parametrize("sem", ["reference", "chained"])
def test_launch_choices_are_bit_identical(ctx, batch, sem):
    shape, dt, dh = batch
    ref_o, ref_c = _fwd(ctx, dt, dh, shape, sem)
    for threads in (64, 128):
        for per in (0, 7, 1):
            ctx.set_tuning(fwd_threads=threads, fwd_layers_per_launch=per)
            o, c = _fwd(ctx, dt, dh, shape, sem)
            assert np.array_equal(o, ref_o) and np.array_equal(c, ref_c), (threads, per)
    for waves in (1, 3, 4):  # register budgets: spills change nothing but speed
        ctx.set_tuning(fwd_threads=256, fwd_layers_per_launch=0, fwd_waves=waves)
        for prec in ("exact", "fast"):
            o, c = _fwd(ctx, dt, dh, shape, sem, prec)
            ctx.set_tuning(fwd_waves=0)
            o0, c0 = _fwd(ctx, dt, dh, shape, sem, prec)
            ctx.set_tuning(fwd_waves=waves)
            assert np.array_equal(o, o0) and np.array_equal(c, c0), (waves, prec)


def test_autotune_forward_picks_a_candidate(ctx, orc, batch):
    shape, dt, dh = batch
    do = torch.empty_like(dt)
    rep = ctx.autotune_forward(dt, dh, do, shape, "reference", reps=2)
    torch.cuda.synchronize()
    want_o, want_c = _oracle_words(orc, dt, dh, shape, "reference")
    # the outputs of the last candidate run are valid results
    assert _equal_words(do.cpu().numpy().view(np.uint32), want_o)
    ents = rep.as_list()
    assert rep.count == len(ents) >= 6 and 0 <= rep.best < rep.count
    assert all(e["ms"] > 0 for e in ents)
    best = ents[rep.best]
    assert best["ms"] == min(e["ms"] for e in ents)
    t = ctx.get_tuning()
    assert (t.fwd_threads, t.fwd_layers_per_launch, t.fwd_waves) == (best["threads"], best["layers_per_launch"],
                                                                      best["waves"])
    assert sorted({e["waves"] for e in ents}) == [0, 1, 3, 4]
    o, c = _fwd(ctx, dt, dh, shape, "reference")
    assert _equal_words(o, want_o) and _equal_words(c, want_c)  # the chosen launch, against the oracle
    ctx.set_tuning(fwd_threads=256, fwd_layers_per_launch=0, fwd_waves=0)
    o0, c0 = _fwd(ctx, dt, dh, shape, "reference")
    assert np.array_equal(o, o0) and np.array_equal(c, c0)


def test_repack_round_trip_and_width_invariance(ctx, batch):
    """32 -> 16 -> 32 restores every word; the forward on the width-16 copy equals the width-32 forward
    after repacking its output back."""
    shape, dt, dh = batch
    t16, h16 = torch.empty_like(dt), torch.empty_like(dh)
    ctx.repack(dt, 32, t16, 16, shape.nevts, shape.ntrks, 1, 28)
    ctx.repack(dh, 32, h16, 16, shape.nevts, shape.ntrks, shape.nlayer, 9)
    t32, h32 = torch.empty_like(dt), torch.empty_like(dh)
    ctx.repack(t16, 16, t32, 32, shape.nevts, shape.ntrks, 1, 28)
    ctx.repack(h16, 16, h32, 32, shape.nevts, shape.ntrks, shape.nlayer, 9)
    torch.cuda.synchronize()
    assert torch.equal(t32.view(torch.int32), dt.view(torch.int32))
    assert torch.equal(h32.view(torch.int32), dh.view(torch.int32))
    # the layout itself: track t of event e, field f
    T, T16 = dt.cpu().numpy().reshape(-1, 28, 32), t16.cpu().numpy().reshape(-1, 28, 16)
    for e, t, f in ((0, 0, 0), (1, 37, 5), (2, 1023, 27)):
        a = T[e * 32 + t // 32, f, t % 32].view(np.uint32)
        b = T16[e * 64 + t // 16, f, t % 16].view(np.uint32)
        assert a == b
    s16 = Shape(shape.nevts, shape.ntrks, shape.nlayer, 16)
    for sem in ("reference", "chained"):
        o32, c32 = _fwd(ctx, dt, dh, shape, sem)
        o16 = torch.empty_like(dt)
        c16 = torch.empty(shape.ntrk_total, device="cuda")
        ctx.forward(t16, h16, o16, s16, sem, chi2=c16, precision="exact")
        back = torch.empty_like(dt)
        ctx.repack(o16, 16, back, 32, shape.nevts, shape.ntrks, 1, 28)
        torch.cuda.synchronize()
        assert np.array_equal(back.cpu().numpy().view(np.uint32), o32)
        assert np.array_equal(c16.cpu().numpy().view(np.uint32), c32)  # chi2 is indexed by flat track


def test_autotune_width_reports_each_width(ctx, batch):
    shape, dt, dh = batch
    rep = ctx.autotune_width(dt, dh, shape, widths=(16, 32, 64), reps=2)
    ents = rep.as_list()
    assert [e["bsize"] for e in ents] == [16, 32, 64] and all(e["ms"] > 0 for e in ents)


def test_autotune_ragged_and_adjoint(ctx):
    import p2r_amd
    counts = p2r_amd.lognormal_counts(12, median=600, sigma=0.5, lo=30, hi=3000, seed=5)
    trk, hits, mask, offs = p2r_amd.gen_ragged(counts, 20, 32, kmin=5, kmax=20)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    dm = torch.from_numpy(mask.view(np.int32)).cuda()
    do = torch.empty_like(dt)
    ctx.forward_ragged(dt, dh, do, offs, dm, 20, 32, "chained")
    torch.cuda.synchronize()
    ref = do.cpu().numpy().view(np.uint32).copy()
    rep = ctx.autotune_ragged(dt, dh, do, offs, dm, 20, 32, "chained", reps=2)
    ents = rep.as_list()
    assert rep.count == 7 and ctx.get_tuning().ragged_schedule == ents[rep.best]["schedule"]
    assert -2 in [e["schedule"] for e in ents]
    for threads in (64, 128):
        ctx.set_tuning(ragged_schedule=-1, ragged_threads=threads)
        ctx.forward_ragged(dt, dh, do, offs, dm, 20, 32, "chained")
        torch.cuda.synchronize()
        assert np.array_equal(do.cpu().numpy().view(np.uint32), ref), threads
    shape = Shape(1, 2048, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    do = torch.empty_like(dt)
    rep = ctx.autotune_adjoint(dt, dh, do, shape, grad_precision="f32", reps=1)
    ents = rep.as_list()
    assert [e["checkpoint_stride"] for e in ents] == [1, 2, 4, 8]
    assert ctx.get_tuning().checkpoint_stride == ents[rep.best]["checkpoint_stride"]


@pytest.mark.parametrize("nstreams", [1, 3, 5])
def test_streamed_host_to_host_equals_forward(ctx, orc, batch, nstreams):
    """p2r_forward_streamed (chunks of blocks over streams, pinned host buffers in and out) returns the
    oracle's words, as p2r_forward on device buffers does; 96 blocks over 5 streams leaves uneven chunks."""
    shape, dt, dh = batch
    for sem in ("reference", "chained"):
        want, _ = _oracle_words(orc, dt, dh, shape, sem)
        dev, _ = _fwd(ctx, dt, dh, shape, sem)
        assert _equal_words(dev, want), sem
        ht, hh = dt.cpu().pin_memory(), dh.cpu().pin_memory()
        ho = torch.zeros_like(ht).pin_memory()
        d_in, d_hit, d_out = torch.empty_like(dt), torch.empty_like(dh), torch.empty_like(dt)
        ctx.forward_streamed(ht, hh, ho, d_in, d_hit, d_out, shape, sem, nstreams=nstreams)
        assert _equal_words(ho.numpy().view(np.uint32), want), sem
