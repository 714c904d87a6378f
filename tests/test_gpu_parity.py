"""GPU parity: the HIP path (through the C-ABI) against the oracle on the same inputs.

P2R_EXACT arithmetic must be BIT-IDENTICAL to the oracle's IEEE-strict FP32
restatement run with correctly rounded sin/cos (the oracle with libm's sinf/cosf
is itself bit-identical to the reference's compiled routines; the two oracle
modes differ only in those four trig values per layer).  NaN outputs compare
equal to NaN regardless of payload.

P2R_FAST arithmetic (fused multiply-adds, hardware reciprocal) is held to
parity.compare_statistical: at least as close to the IEEE-strict CPU result as
the reference's own compiled CPU build (build.py flags) is, and no less
accurate than the CPU path against the FP64 replay, field by field.
"""
import json

import numpy as np
import pytest

import oracle
import parity
from p2r_amd import Shape

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    import p2r_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return p2r_amd.Context(0)


@pytest.fixture(scope="module")
def orc_alt():
    return oracle.Oracle("liboracle_bench.so")


def run_gpu(ctx, trk, hits, shape, semantics, precision="exact", chi2=False):
    dt = torch.from_numpy(trk).cuda()
    dh = torch.from_numpy(hits).cuda()
    do = torch.empty_like(dt)
    dc = torch.full((shape.ntrk_total,), float("nan"), dtype=torch.float32, device="cuda") if chi2 else None
    ctx.forward(dt, dh, do, shape, semantics, chi2=dc, precision=precision)
    torch.cuda.synchronize()
    return do.cpu().numpy(), (dc.cpu().numpy() if chi2 else None)


def bit_mismatches(a, b):
    """Number of float elements that differ (NaN == NaN)."""
    a = np.asarray(a, np.float32).ravel()
    b = np.asarray(b, np.float32).ravel()
    same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
    return int((~same).sum())


def check_exact(orc, ctx, trk, hits, shape, semantics, label):
    sem = 0 if semantics == "reference" else 1
    print(f"[{label}] exact shape={shape} semantics={semantics}", flush=True)
    gpu, c2g = run_gpu(ctx, trk, hits, shape, semantics, "exact", chi2=True)
    cpu, c2c = orc.run_f32(trk, hits, shape.nblk, shape.nlayer, shape.bsize, semantics=sem, chi2=True, cr_trig=True)
    g = gpu.reshape(-1, 28, shape.bsize)
    c = cpu.reshape(-1, 28, shape.bsize)
    n_state = bit_mismatches(g[:, :27], c[:, :27])
    n_q = int((g[:, 27].view(np.int32) != c[:, 27].view(np.int32)).sum())
    n_chi2 = bit_mismatches(c2g, c2c)
    print(f"[{label}] mismatching state elements={n_state} charges={n_q} chi2={n_chi2} "
          f"of {shape.ntrk_total} tracks", flush=True)
    assert n_state == 0 and n_q == 0 and n_chi2 == 0
    return gpu, cpu


def check_fast(orc, orc_alt, ctx, trk, hits, shape, semantics, label):
    sem = 0 if semantics == "reference" else 1
    print(f"[{label}] fast shape={shape} semantics={semantics}", flush=True)
    gpu, _ = run_gpu(ctx, trk, hits, shape, semantics, "fast")
    cpu = orc.run_f32(trk, hits, shape.nblk, shape.nlayer, shape.bsize, semantics=sem)
    alt = orc_alt.run_f32(trk, hits, shape.nblk, shape.nlayer, shape.bsize, semantics=sem)
    par64, cov64, _ = orc.run_f64(trk, hits, shape.nblk, shape.nlayer, shape.bsize, semantics=sem)
    res = parity.compare_statistical(gpu, cpu, alt, shape.bsize, par64, cov64)
    print(label, json.dumps(res), flush=True)
    assert res["ok"], res
    return res


# ---- P2R_EXACT: bit-identical --------------------------------------------------

@pytest.mark.parametrize("semantics", ["reference", "chained"])
def test_exact_reference_inputs(orc, ctx, semantics):
    """C0: the reference's own generated inputs, 10 x 9600 x 20, bsize 32."""
    import p2r_amd
    shape = Shape(10, 9600, 20, 32)
    trk, hits = p2r_amd.gen_reference(shape)
    gpu, _ = check_exact(orc, ctx, trk, hits, shape, semantics, "c0")
    if semantics == "reference":
        # the printed summary of the GPU output vs the reference program's own output lines
        st, cn = orc.summary(gpu, shape.nevts, shape.ntrks, shape.bsize, orc.c0_hit(19))
        st_c, cn_c = orc.summary(orc.run_f32(trk, hits, shape.nblk, 20, 32), shape.nevts, shape.ntrks, 32,
                                 orc.c0_hit(19))
        lines = oracle.summary_lines(st, cn)
        print("\n".join(lines))
        assert list(cn) == list(cn_c)
        np.testing.assert_allclose(st, st_c, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("semantics", ["reference", "chained"])
def test_exact_10x9600x20(orc, ctx, semantics, mode):
    """10 events x 9600 tracks x 20 layers of seeded benchmark inputs, consistent (0) and literal (1) inputs."""
    import p2r_amd
    shape = Shape(10, 9600, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape, p2r_amd.gen_config(mode=mode))
    check_exact(orc, ctx, trk, hits, shape, semantics, f"10x9600x20-mode{mode}")


@pytest.mark.parametrize("bsize", [1, 16, 64, 128])
def test_exact_bsize_variants(orc, ctx, bsize):
    import p2r_amd
    shape = Shape(2, 1024 if bsize != 1 else 256, 20, bsize)
    trk, hits = p2r_amd.gen_uniform(shape)
    for sem in ("reference", "chained"):
        check_exact(orc, ctx, trk, hits, shape, sem, f"bsize{bsize}")


def test_exact_nonfinite_inputs(orc, ctx):
    import p2r_amd
    shape = Shape(1, 256, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    t = trk.reshape(shape.nblk, 28, 32)
    t[0, 0, 3] = np.nan      # x of track 3
    t[1, 9, 5] = np.inf      # a covariance entry of track 37
    t[2, 3, 7] = 0.0         # 1/pt = 0 of track 71
    h = hits.reshape(shape.nblk, 20, 9, 32)
    h[3, 19, 2, 11] = np.nan  # last-layer hit z of track 107
    for sem in ("reference", "chained"):
        gpu, _ = check_exact(orc, ctx, trk, hits, shape, sem, "nonfinite")
        assert not np.isfinite(gpu.reshape(-1, 28, 32)[0, :27, 3]).all()


def test_exact_reciprocal_range_edges(orc, ctx):
    """Finite inputs at the edges of the reciprocal's range, one lane each in otherwise ordinary wavefronts:
    1/pt = 2^126 (pt = 2^-126, the smallest normal), 1/pt = 3e38 (pt denormal), 1/pt denormal (pt overflows)
    and 1/pt = 2^125.  Every word equals the oracle's, with and without chi2."""
    import p2r_amd
    shape = Shape(1, 256, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    t = trk.reshape(shape.nblk, 28, 32)
    t[0, 3, 5] = np.float32(2.0 ** 126)       # pt = 2^-126: the largest input outside the verified range
    t[1, 3, 9] = np.float32(3.0e38)           # pt denormal
    t[2, 3, 1] = np.float32(2.0e-39)          # 1/pt denormal -> pt overflows
    t[3, 3, 30] = np.float32(2.0 ** 125)      # just inside the range
    for sem in ("reference", "chained"):
        _, cpu = check_exact(orc, ctx, trk, hits, shape, sem, "rcp range")
        gpu, _ = run_gpu(ctx, trk, hits, shape, sem, "exact", chi2=False)  # the instance without chi2 as well
        assert bit_mismatches(gpu, cpu) == 0


def test_exact_block_entry_points(orc, ctx):
    """p2r_propagateToR / p2r_KalmanUpdate on MPTRK/MPHIT arrays vs the oracle's block routines."""
    import p2r_amd
    bs = 32
    shape = Shape(1, 256, 3, bs)
    trk, hits = p2r_amd.gen_uniform(shape)
    nblk = shape.nblk
    dt = torch.from_numpy(trk).cuda()
    dh = torch.from_numpy(hits).cuda()
    do = torch.zeros_like(dt)
    kf = p2r_amd.default_params().kfact
    lay = 1
    ctx.propagateToR(dt[6 * bs:], dt[0:], dt[27 * bs:], dh[lay * 9 * bs:], do[6 * bs:], do[0:], bs, nblk,
                     28 * bs, shape.nlayer * 9 * bs, kfact=kf)
    ctx.KalmanUpdate(do[6 * bs:], do[0:], dh[lay * 9 * bs + 3 * bs:], dh[lay * 9 * bs:], bs, nblk, 28 * bs,
                     shape.nlayer * 9 * bs)
    torch.cuda.synchronize()
    gpu = do.cpu().numpy().reshape(nblk, 28, bs)
    orc.lib.orc_set_trig_mode(1)
    try:
        for b in range(nblk):
            tb = trk.reshape(nblk, 28 * bs)[b]
            hb = hits.reshape(nblk, shape.nlayer, 9 * bs)[b, lay]
            e, p = orc.propagate_block(bs, kf, tb[6 * bs:27 * bs].copy(), tb[:6 * bs].copy(),
                                       tb[27 * bs:].view(np.int32).copy(), hb[:3 * bs].copy())
            e, p = orc.kalman_block(bs, e, p, hb[3 * bs:].copy(), hb[:3 * bs].copy())
            assert bit_mismatches(gpu[b, :27].ravel(), np.concatenate([p, e])) == 0
    finally:
        orc.lib.orc_set_trig_mode(0)


# ---- P2R_FAST: FP32 parity criterion ------------------------------------------

@pytest.mark.parametrize("semantics", ["reference", "chained"])
def test_fast_reference_inputs(orc, orc_alt, ctx, semantics):
    import p2r_amd
    shape = Shape(10, 9600, 20, 32)
    trk, hits = p2r_amd.gen_reference(shape)
    check_fast(orc, orc_alt, ctx, trk, hits, shape, semantics, "c0")


@pytest.mark.parametrize("semantics,mode", [("reference", 0), ("chained", 0), ("reference", 1)])
def test_fast_10x9600x20(orc, orc_alt, ctx, semantics, mode):
    """Consistent inputs in both semantics; literal inputs in reference semantics only (literal chained
    fits break down in FP32 for ~90 % of tracks; they are covered bit-exactly by the exact tests)."""
    import p2r_amd
    shape = Shape(10, 9600, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape, p2r_amd.gen_config(mode=mode))
    check_fast(orc, orc_alt, ctx, trk, hits, shape, semantics, f"10x9600x20-mode{mode}")


@pytest.mark.parametrize("semantics", ["reference", "chained"])
def test_fast_full_headline_shape(orc, orc_alt, ctx, semantics):
    """The fast path on the whole 100 x 9600 x 20 batch (the benchmark's inputs), by the same statistical
    criterion as the config-1 tests: no worse against the strict oracle than the reference's own fast CPU
    build is, and close to the FP64 replay."""
    import p2r_amd
    shape = Shape(100, 9600, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    check_fast(orc, orc_alt, ctx, trk, hits, shape, semantics, "full")


# ---- errors and work accounting ---------------------------------------------

def test_shape_errors(ctx):
    import p2r_amd
    t = torch.zeros(28 * 32, device="cuda")
    h = torch.zeros(9 * 32, device="cuda")
    with pytest.raises(p2r_amd.P2RError) as e:
        ctx.forward(t, h, t, Shape(1, 33, 1, 32))
    assert e.value.status == p2r_amd.P2R_ERR_SHAPE
    with pytest.raises(p2r_amd.P2RError) as e:
        ctx.forward(t, h, t, Shape(1, 32, 1, 32), semantics=7)
    assert e.value.status == p2r_amd.P2R_ERR_ARG
    ctx.forward(t, h, t, Shape(0, 32, 1, 32))  # empty batch is a no-op


@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_work_scales_with_layers(ctx, precision):
    """Anti-DCE: time(nlayer=20) / time(nlayer=10) ~ 2 in both semantics (every layer is computed).  The
    two shapes are timed in alternation after the GPU's clocks have ramped up, best of 3 rounds each, at a
    size where a launch takes 0.15-0.6 ms (launch overhead and clock ramp-up do not enter the ratio)."""
    import time

    import p2r_amd
    for sem in ("reference", "chained"):
        runs = {}
        for nl in (10, 20):
            shape = Shape(100, 9600, nl, 32)
            trk, hits = p2r_amd.gen_uniform(shape)
            dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
            runs[nl] = (shape, dt, dh, torch.empty_like(dt))
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 0.3:  # clock ramp-up
            for shape, dt, dh, do in runs.values():
                ctx.forward(dt, dh, do, shape, sem, precision=precision)
            torch.cuda.synchronize()
        times = {10: float("inf"), 20: float("inf")}
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(3):
            for nl, (shape, dt, dh, do) in runs.items():
                e0.record()
                for _ in range(10):
                    ctx.forward(dt, dh, do, shape, sem, precision=precision)
                e1.record()
                torch.cuda.synchronize()
                times[nl] = min(times[nl], e0.elapsed_time(e1) / 10)
        ratio = times[20] / times[10]
        print(f"work scaling {precision} {sem}: t10={times[10]:.4f} ms t20={times[20]:.4f} ms ratio={ratio:.3f}",
              flush=True)
        assert 1.6 < ratio < 2.4, (sem, times)


# ---- uneven batch (persistent ragged kernel) --------------------------------

def _ragged_inputs(nevts=40, median=600, bsize=32, nlayer=20, seed=11):
    import p2r_amd
    counts = p2r_amd.lognormal_counts(nevts, median=median, sigma=0.5, lo=30, hi=3000, seed=seed)
    trk, hits, mask, offs = p2r_amd.gen_ragged(counts, nlayer, bsize, kmin=min(5, nlayer), kmax=min(20, nlayer))
    return counts, trk, hits, mask, offs


def run_ragged(ctx, trk, hits, mask, offs, nlayer, bsize, semantics, precision="exact", rounds=0):
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    dm = torch.from_numpy(mask.view(np.int32)).cuda()
    do = torch.empty_like(dt)
    dc = torch.full((mask.size,), float("nan"), dtype=torch.float32, device="cuda")
    ctx.forward_ragged(dt, dh, do, offs, dm, nlayer, bsize, semantics, chi2=dc, precision=precision,
                       tile_rounds=rounds)
    torch.cuda.synchronize()
    return do.cpu().numpy(), dc.cpu().numpy()


@pytest.mark.parametrize("rounds", [-1, -2, 1, 2, 4])
@pytest.mark.parametrize("semantics", ["chained", "reference"])
def test_exact_ragged_matches_oracle(orc, ctx, semantics, rounds):
    """Uneven events + masked layers, one persistent launch: bit-identical to the oracle's masked run."""
    counts, trk, hits, mask, offs = _ragged_inputs()
    nblk = int(offs[-1])
    sem = 0 if semantics == "reference" else 1
    gpu, c2g = run_ragged(ctx, trk, hits, mask, offs, 20, 32, semantics, rounds=rounds)
    cpu, c2c = orc.run_f32(trk, hits, nblk, 20, 32, semantics=sem, chi2=True, mask=mask, cr_trig=True)
    n = bit_mismatches(gpu.reshape(-1, 28, 32)[:, :27], cpu.reshape(-1, 28, 32)[:, :27])
    nq = int((gpu.reshape(-1, 28, 32)[:, 27].view(np.int32) != cpu.reshape(-1, 28, 32)[:, 27].view(np.int32)).sum())
    nc = bit_mismatches(c2g, c2c)
    print(f"ragged {semantics} rounds={rounds}: {int(counts.sum())} tracks in {nblk} blocks, mismatches "
          f"state={n} q={nq} chi2={nc}", flush=True)
    assert n == 0 and nq == 0 and nc == 0


@pytest.mark.parametrize("rounds", [-1, -2, 1])
@pytest.mark.parametrize("semantics", ["chained", "reference"])
def test_fast_ragged_meets_the_fast_criterion(orc, orc_alt, ctx, semantics, rounds):
    """Uneven batch with fast arithmetic, every schedule: each kernel is its own compilation of the fused
    formulas, so the schedules are held to the fast-mode criterion against the masked oracle run (as close
    to the IEEE-strict result as the reference's own fast-math build), not to each other's bits."""
    counts, trk, hits, mask, offs = _ragged_inputs()
    nblk = int(offs[-1])
    sem = 0 if semantics == "reference" else 1
    gpu, _ = run_ragged(ctx, trk, hits, mask, offs, 20, 32, semantics, precision="fast", rounds=rounds)
    cpu = orc.run_f32(trk, hits, nblk, 20, 32, semantics=sem, mask=mask)
    alt = orc_alt.run_f32(trk, hits, nblk, 20, 32, semantics=sem, mask=mask)
    par64, cov64, _ = orc.run_f64(trk, hits, nblk, 20, 32, semantics=sem, mask=mask)
    res = parity.compare_statistical(gpu, cpu, alt, 32, par64, cov64)
    print(f"fast ragged {semantics} rounds={rounds}", json.dumps(res), flush=True)
    assert res["ok"], res


@pytest.mark.parametrize("rounds", [-1, -2, 1])
@pytest.mark.parametrize("bsize", [1, 8, 64])
def test_exact_ragged_other_widths(orc, ctx, bsize, rounds):
    """Matriplex widths other than 32 through both schedules (chained, chi2 on)."""
    counts, trk, hits, mask, offs = _ragged_inputs(nevts=12, bsize=bsize)
    nblk = int(offs[-1])
    gpu, c2g = run_ragged(ctx, trk, hits, mask, offs, 20, bsize, "chained", rounds=rounds)
    cpu, c2c = orc.run_f32(trk, hits, nblk, 20, bsize, semantics=1, chi2=True, mask=mask, cr_trig=True)
    n = bit_mismatches(gpu.reshape(-1, 28, bsize)[:, :27], cpu.reshape(-1, 28, bsize)[:, :27])
    assert n == 0 and bit_mismatches(c2g, c2c) == 0


@pytest.mark.parametrize("rounds", [-1, -2])
@pytest.mark.parametrize("semantics", ["chained", "reference"])
def test_exact_ragged_widest_matriplex(orc, ctx, semantics, rounds):
    """bsize 1024 (the largest width): a pair-stream tile lies inside one or two blocks, the hit offsets and
    field strides are at their largest, and events end mid-block."""
    import p2r_amd
    counts = np.array([1500, 0, 2100, 700], np.int32)
    trk, hits, mask, offs = p2r_amd.gen_ragged(counts, 20, 1024, kmin=5, kmax=20)
    nblk = int(offs[-1])
    gpu, c2g = run_ragged(ctx, trk, hits, mask, offs, 20, 1024, semantics, rounds=rounds)
    cpu, c2c = orc.run_f32(trk, hits, nblk, 20, 1024, semantics=0 if semantics == "reference" else 1, chi2=True,
                           mask=mask, cr_trig=True)
    assert bit_mismatches(gpu, cpu) == 0 and bit_mismatches(c2g, c2c) == 0


def test_ragged_full_masks_equal_uniform(ctx):
    """All layers present and equal event sizes: the ragged launch reproduces p2r_forward bit for bit."""
    import p2r_amd
    shape = Shape(6, 640, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    offs = np.arange(shape.nevts + 1, dtype=np.int64) * shape.nb
    mask = np.full(shape.ntrk_total, (1 << 20) - 1, np.uint32)
    for sem in ("reference", "chained"):
        uni, _ = run_gpu(ctx, trk, hits, shape, sem)
        for rounds in (-1, -2):
            rag, _ = run_ragged(ctx, trk, hits, mask, offs, 20, 32, sem, rounds=rounds)
            assert bit_mismatches(uni, rag) == 0, (sem, rounds)


@pytest.mark.parametrize("rounds", [-1, -2, 1])
def test_ragged_ignores_mask_bits_beyond_nlayer(orc, ctx, rounds):
    counts, trk, hits, mask, offs = _ragged_inputs(nevts=8, nlayer=12)
    junk = mask | np.uint32(0xFFFFF000)  # bits 12..31 name no hit block
    gpu, _ = run_ragged(ctx, trk, hits, junk, offs, 12, 32, "chained", rounds=rounds)
    cpu = orc.run_f32(trk, hits, int(offs[-1]), 12, 32, semantics=1, mask=mask, cr_trig=True)
    assert bit_mismatches(gpu.reshape(-1, 28, 32)[:, :27], cpu.reshape(-1, 28, 32)[:, :27]) == 0


@pytest.mark.parametrize("nlayer,density", [(1, 0.5), (5, 0.1), (20, 0.03), (20, 0.97), (32, 0.5), (32, 0.0)])
@pytest.mark.parametrize("semantics", ["chained", "reference"])
def test_exact_pair_stream_mask_densities(orc, ctx, semantics, nlayer, density):
    """The pair-stream schedule over masks from empty to full (windows with no pair, steps short of 64 pairs,
    steps spanning two layers), 32-bit masks, and a tile cut by the end of the batch (13 blocks of 8 slots)."""
    counts = np.array([40, 17, 0, 33], np.int32)
    import p2r_amd
    trk, hits, _, offs = p2r_amd.gen_ragged(counts, nlayer, 8, kmin=1, kmax=nlayer)
    rng = np.random.default_rng(nlayer * 100 + int(density * 100))
    bits = rng.random((offs[-1] * 8, nlayer)) < density
    mask = (bits * (1 << np.arange(nlayer, dtype=np.uint64))).sum(1).astype(np.uint32)
    sem = 0 if semantics == "reference" else 1
    gpu, c2g = run_ragged(ctx, trk, hits, mask, offs, nlayer, 8, semantics, rounds=-2)
    cpu, c2c = orc.run_f32(trk, hits, int(offs[-1]), nlayer, 8, semantics=sem, chi2=True, mask=mask, cr_trig=True)
    assert bit_mismatches(gpu, cpu) == 0 and bit_mismatches(c2g, c2c) == 0


def test_ragged_graph_replays_match_eager(ctx):
    """The persistent launch captured in a HIP graph: every replay (with eager work in between) equals the
    eager result bit for bit."""
    counts, trk, hits, mask, offs = _ragged_inputs(nevts=16)
    eager, _ = run_ragged(ctx, trk, hits, mask, offs, 20, 32, "chained")
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    dm = torch.from_numpy(mask.view(np.int32)).cuda()
    do = torch.zeros_like(dt)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx.forward_ragged(dt, dh, do, offs, dm, 20, 32, "chained", stream=s)  # warm-up outside capture
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            ctx.forward_ragged(dt, dh, do, offs, dm, 20, 32, "chained", stream=s)
    s.synchronize()
    for i in range(3):
        do.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert bit_mismatches(do.cpu().numpy(), eager) == 0, f"replay {i}"
        _ = (dt * 2).sum().item()  # unrelated eager work between replays


# ---- options: phi wrap, charge from the sign of 1/pt -------------------------

@pytest.mark.parametrize("options", [1, 2, 3])
@pytest.mark.parametrize("semantics", ["reference", "chained"])
def test_exact_options_match_oracle(orc, ctx, semantics, options):
    """Inputs with phi moved by multiples of 2 pi and tiny 1/pt on some tracks: the options act, and the exact
    path stays bit-identical to the oracle run with the same options (charge word included)."""
    import p2r_amd
    from test_oracle_pins import _options_inputs
    shape, trk, hits = _options_inputs(seed=options)
    sem = 0 if semantics == "reference" else 1
    p = p2r_amd.default_params("exact", options=options)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    do = torch.empty_like(dt)
    dc = torch.empty(shape.ntrk_total, device="cuda")
    ctx.forward(dt, dh, do, shape, semantics, params=p, chi2=dc)
    torch.cuda.synchronize()
    cpu, c2c = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=sem, chi2=True, cr_trig=True, options=options)
    gpu = do.cpu().numpy()
    assert bit_mismatches(gpu, cpu) == 0 and bit_mismatches(dc.cpu().numpy(), c2c) == 0
    base = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=sem, cr_trig=True)
    assert bit_mismatches(base, cpu) > 0  # the options changed something on these inputs


@pytest.mark.parametrize("rounds", [-1, -2, 1])
def test_exact_options_ragged_match_oracle(orc, ctx, rounds):
    import p2r_amd
    counts, trk, hits, mask, offs = _ragged_inputs(nevts=10)
    T = trk.reshape(-1, 28, 32)
    T[:, 4, :] += np.float32(2 * np.pi)
    p = p2r_amd.default_params("exact", options=3)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    dm = torch.from_numpy(mask.view(np.int32)).cuda()
    do = torch.empty_like(dt)
    ctx.forward_ragged(dt, dh, do, offs, dm, 20, 32, "chained", params=p, tile_rounds=rounds)
    torch.cuda.synchronize()
    cpu = orc.run_f32(trk, hits, int(offs[-1]), 20, 32, semantics=1, mask=mask, cr_trig=True, options=3)
    assert bit_mismatches(do.cpu().numpy(), cpu) == 0


def test_options_rejected_where_unsupported(ctx):
    import p2r_amd
    shape = Shape(1, 64, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    with pytest.raises(p2r_amd.P2RError, match="not supported"):
        ctx.forward_adjoint(dt, dh, torch.empty_like(dt), shape, params=p2r_amd.default_params("exact", options=1))
    with pytest.raises(p2r_amd.P2RError):
        ctx.forward(dt, dh, torch.empty_like(dt), shape, params=p2r_amd.default_params("exact", options=8))


@pytest.mark.parametrize("semantics", ["reference", "chained"])
def test_exact_full_headline_shape_matches_oracle(orc, ctx, semantics):
    """The whole 100 events x 9600 tracks x 20 layers batch (19.2M updates) through the exact path with the
    tuned register budget, every word compared with the oracle (run threaded on the host)."""
    import p2r_amd
    shape = Shape(100, 9600, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    ctx.set_tuning(fwd_waves=3)
    try:
        check_exact(orc, ctx, trk, hits, shape, semantics, "full")
    finally:
        ctx.set_tuning(fwd_waves=0)


# ---- chi2 per layer (P2R_OPT_CHI2_PER_LAYER) ---------------------------------

@pytest.mark.parametrize("per", [0, 7])
@pytest.mark.parametrize("semantics", ["reference", "chained"])
def test_exact_chi2_per_layer_matches_oracle(orc, ctx, semantics, per):
    """Every update's chi2, [layer][track], bit-identical to the oracle; the state output is unchanged by the
    option; also with the layer loop split over launches."""
    import p2r_amd
    shape = Shape(2, 640, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    sem = 0 if semantics == "reference" else 1
    n = shape.ntrk_total
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    do = torch.empty_like(dt)
    dc = torch.full((20 * n,), float("nan"), device="cuda")
    ctx.set_tuning(fwd_layers_per_launch=per)
    try:
        ctx.forward(dt, dh, do, shape, semantics, params=p2r_amd.default_params("exact", options=4), chi2=dc)
        torch.cuda.synchronize()
    finally:
        ctx.set_tuning(fwd_layers_per_launch=0)
    cpu, c2 = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=sem, chi2=True, cr_trig=True, options=4)
    assert bit_mismatches(do.cpu().numpy(), cpu) == 0
    assert bit_mismatches(dc.cpu().numpy(), c2) == 0


@pytest.mark.parametrize("rounds", [-1, -2, 1])
@pytest.mark.parametrize("semantics", ["chained", "reference"])
def test_exact_ragged_chi2_per_layer_matches_oracle(orc, ctx, semantics, rounds):
    """Uneven batch with chi2 per layer: present layers bit-identical to the oracle, absent layers 0."""
    import p2r_amd
    counts, trk, hits, mask, offs = _ragged_inputs(nevts=10)
    nblk, nslot = int(offs[-1]), mask.size
    sem = 0 if semantics == "reference" else 1
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    dm = torch.from_numpy(mask.view(np.int32)).cuda()
    do = torch.empty_like(dt)
    dc = torch.full((20 * nslot,), float("nan"), device="cuda")
    ctx.forward_ragged(dt, dh, do, offs, dm, 20, 32, semantics, params=p2r_amd.default_params("exact", options=4),
                       chi2=dc, tile_rounds=rounds)
    torch.cuda.synchronize()
    cpu, c2 = orc.run_f32(trk, hits, nblk, 20, 32, semantics=sem, chi2=True, mask=mask, cr_trig=True, options=4)
    assert bit_mismatches(do.cpu().numpy(), cpu) == 0
    assert bit_mismatches(dc.cpu().numpy(), c2) == 0
