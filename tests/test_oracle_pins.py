"""Pins of the oracle (CPU restatement of propagate-tor-test_tbb.cpp).

Each pin is chosen so a plausible transcription mistake (dropped term, wrong
sign or index, transposed operand) fails at least one of them:
  * bit-exact against the reference's own compiled routines (oracle/_ref/,
    built from src/propagate-tor-test_acc.v3.cpp and _pstl.cpp) and against
    golden vectors captured from them (tests/golden/),
  * the reference program's printed summary at 10 x 9600 x 20,
  * closed forms: glibc rand() prefix, kfact, textbook Kalman update (FP64),
    covariance transport = J P J^T, layer-semantics invariants.
"""
import json
import os

import numpy as np
import pytest

from golden.make_ref_vectors import BS, wide_block

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _run_oracle_block(orc, kfact, par, cv, q, hits):
    outs = []
    for hp, hc in hits:
        e, p = orc.propagate_block(BS, kfact, cv.ravel(), par.ravel(), q, hp.ravel())
        ke, kp = orc.kalman_block(BS, e, p, hc.ravel(), hp.ravel())
        outs.append(np.concatenate([p, e, kp, ke]))
    return np.stack(outs)


def _c0_block_inputs(orc, nlayer=8):
    trk, hits = orc.gen_c0(2, 64, 20, BS)
    blk = trk[:28 * BS]
    hl = []
    for l in range(nlayer):
        hb = hits[l * 9 * BS:(l + 1) * 9 * BS]
        hl.append((hb[:3 * BS].reshape(3, BS), hb[3 * BS:].reshape(6, BS)))
    return blk[:6 * BS].reshape(6, BS), blk[6 * BS:27 * BS].reshape(21, BS), blk[27 * BS:].view(np.int32).copy(), hl


def test_glibc_rand_prefix(orc):
    # glibc's TYPE_3 additive generator after srand(1) (the reference never seeds)
    assert list(orc.rand_prefix(1, 5)) == [1804289383, 846930886, 1681692777, 1714636915, 1957747793]


def test_kfact_closed_form(orc):
    # tbb.cpp:654: 100/(-0.299792458*3.8112) evaluated in double, narrowed to float
    assert orc.kfact_tbb() == np.float32(100 / (-0.299792458 * 3.8112))
    assert abs(orc.kfact_tbb() + 87.52206) < 1e-4


def test_block_routines_match_golden_vectors(orc):
    g = np.load(os.path.join(GOLDEN, "ref_block_vectors.npz"))
    kfact = float(g["kfact"])
    got_c0 = _run_oracle_block(orc, kfact, *_c0_block_inputs(orc))
    assert np.array_equal(_bits(got_c0), _bits(g["c0"]))
    got_w = _run_oracle_block(orc, kfact, *wide_block(int(g["wide_seed"]), 4))
    assert np.array_equal(_bits(got_w), _bits(g["wide"]))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_block_routines_match_reference_live(orc, ref, seed):
    par, cv, q, hits = wide_block(seed, 6)
    want = _run_oracle_block(orc, ref.kfact, par, cv, q, hits)
    from golden.make_ref_vectors import run_ref_block
    got = run_ref_block(ref, par, cv, q, hits)
    assert np.array_equal(_bits(want), _bits(got))


def test_generator_c0_matches_golden(orc):
    g = np.load(os.path.join(GOLDEN, "ref_gen_c0_head.npz"))
    nevts, ntrks, nlayer, bs = (int(v) for v in g["dims"])
    trk, hits = orc.gen_c0(nevts, ntrks, nlayer, bs)
    assert np.array_equal(_bits(trk[:2 * 28 * bs]), _bits(g["trk_blocks"]))
    assert np.array_equal(_bits(hits[:2 * 9 * bs]), _bits(g["hits_layer0"]))


def test_generator_c0_matches_reference_live(orc, ref):
    dims, trk_r, hits_r = ref.pstl_generate()
    trk, hits = orc.gen_c0(*dims)
    assert np.array_equal(_bits(trk), _bits(trk_r))
    assert np.array_equal(_bits(hits), _bits(hits_r))


def test_c0_inputs_table_matches_reference_header(orc):
    d = json.load(open(os.path.join(GOLDEN, "input_track_c0.json")))
    for lay in range(20):
        assert np.array_equal(orc.c0_hit(lay), np.float32(d["hits"][21 - lay]["pos"]))


def test_c0_summary_matches_reference_program(orc):
    import oracle
    trk, hits = orc.gen_c0(10, 9600, 20, BS)
    out = orc.run_f32(trk, hits, 10 * 300, 20, BS, semantics=0)
    st, cn = orc.summary(out, 10, 9600, BS, orc.c0_hit(19))
    want = open(os.path.join(GOLDEN, "c0_summary_ref.txt")).read().split("\n")[:13]
    assert oracle.summary_lines(st, cn) == want


def _sym_full(c):
    P = np.zeros((6, 6))
    k = 0
    for i in range(6):
        for j in range(i + 1):
            P[i, j] = P[j, i] = c[k]
            k += 1
    return P


def _sym_pack(P):
    return np.array([P[i, j] for i in range(6) for j in range(i + 1)])


def _random_pd(rng, n, scale):
    A = rng.standard_normal((n, n))
    return (A @ A.T + n * np.eye(n)) * scale


@pytest.mark.parametrize("seed", range(5))
def test_kalman_update_is_textbook_fp64(orc, seed):
    """K = P H^T S^-1, S = H P H^T + R V R^T, H = R [I3 0], R rows (r00, r01, 0), (0, 0, 1)."""
    rng = np.random.default_rng(seed)
    P = _random_pd(rng, 6, 1e-4)
    Vf = _random_pd(rng, 3, 1e-5)
    par = np.array([-30.0, -20.0, 10.0, 0.3, -2.5, 0.9]) + rng.standard_normal(6) * 0.1
    hit = np.array([-31.0, -19.5, 10.3])
    V = np.array([Vf[0, 0], Vf[1, 0], Vf[1, 1], Vf[2, 0], Vf[2, 1], Vf[2, 2]])
    p_out, c_out, chi2 = orc.kalman_lane_f64(par, _sym_pack(P), hit, V)
    r = np.hypot(hit[0], hit[1])
    Rm = np.array([[-(hit[1] + par[1]) / (2 * r), (hit[0] + par[0]) / (2 * r), 0], [0, 0, 1]])
    H = np.hstack([Rm, np.zeros((2, 3))])
    S = H @ P @ H.T + Rm @ Vf @ Rm.T
    K = P @ H.T @ np.linalg.inv(S)
    res = Rm @ (hit - par[:3])
    np.testing.assert_allclose(p_out, par + K @ res, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(_sym_full(c_out), P - K @ H @ P, rtol=1e-9, atol=1e-16)
    np.testing.assert_allclose(chi2, res @ np.linalg.inv(S) @ res, rtol=1e-10)


@pytest.mark.parametrize("layer", [0, 7, 19])
def test_transport_is_J_P_Jt_fp64(orc, layer):
    d = json.load(open(os.path.join(GOLDEN, "input_track_c0.json")))
    par = np.array(d["track"]["par"])
    cov = np.array(d["track"]["cov"]) * 100
    hit = np.array(d["hits"][21 - layer]["pos"])
    p, c, J = orc.propagate_lane_f64(par, cov, 1, hit, float(orc.kfact_tbb()))
    np.testing.assert_allclose(_sym_full(c), J @ _sym_full(cov) @ J.T, rtol=1e-12, atol=1e-20)
    # structure of errorProp (tbb.cpp:753-807): rows 3 and 5 are unit rows,
    # columns 2 and 5 vanish outside rows 2 and their diagonals
    assert np.array_equal(J[3], np.eye(6)[3]) and np.array_equal(J[5], np.eye(6)[5])
    assert J[0, 2] == J[0, 5] == J[1, 2] == J[1, 5] == J[4, 2] == J[4, 5] == 0.0
    assert J[2, 2] == 1.0
    # the propagated point lies on the hit's cylinder (5 fixed-point steps)
    assert abs(np.hypot(p[0], p[1]) - np.hypot(hit[0], hit[1])) < 1e-6 * np.hypot(hit[0], hit[1])
    # pt, theta unchanged; phi advances by the turning angle
    assert p[3] == par[3] and p[5] == par[5]


def test_reference_semantics_is_last_layer_only(orc):
    """tbb.cpp:894 restarts every layer from the input: output = input updated with the last layer only."""
    trk, hits = orc.gen_c0(1, 64, 20, BS)
    nblk = 2
    full = orc.run_f32(trk, hits, nblk, 20, BS, semantics=0)
    h = hits.reshape(nblk, 20, 9 * BS)
    last = np.ascontiguousarray(h[:, 19:20, :]).ravel()
    one = orc.run_f32(trk, last, nblk, 1, BS, semantics=0)
    assert np.array_equal(_bits(full), _bits(one))


def test_chained_semantics_is_fold_of_single_layers(orc):
    """propagate-tor-test_hip.cu:839-845: layer L starts from layer L-1's updated state."""
    trk, hits = orc.gen_c0(1, 64, 6, BS)
    nblk = 2
    full = orc.run_f32(trk, hits, nblk, 6, BS, semantics=1)
    h = hits.reshape(nblk, 6, 9 * BS)
    cur = trk
    for l in range(6):
        cur = orc.run_f32(cur, np.ascontiguousarray(h[:, l:l + 1, :]).ravel(), nblk, 1, BS, semantics=1)
    assert np.array_equal(_bits(full), _bits(cur))


def test_fp64_replay_close_to_fp32(orc):
    trk, hits = orc.gen_c0(1, 64, 20, BS)
    out = orc.run_f32(trk, hits, 2, 20, BS, semantics=0)
    par64, cov64, _ = orc.run_f64(trk, hits, 2, 20, BS, semantics=0)
    o = out.reshape(2, 28, BS)
    par32 = o[:, :6, :].transpose(0, 2, 1).reshape(-1, 6)
    np.testing.assert_allclose(par32, par64, rtol=1e-5)


def test_opcount_frozen(orc):
    """FLOPs per track-layer update of the reference algorithm (propagateToR + KalmanUpdate)."""
    c = orc.opcount()
    assert c == {"add": 494, "mul": 757, "div": 12, "sqrt": 8, "trig": 4, "trivial": 63, "chi2": 9}
    total = c["add"] + c["mul"] + c["div"] + c["sqrt"] + c["trig"]
    assert total == 1275 and total - c["trivial"] == 1212


def _options_inputs(seed=3):
    """A uniform batch with phi moved by multiples of 2 pi on some tracks (the wrap then acts) and 1/pt
    shrunk to ~0 on others (an update can then make it negative)."""
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(1, 256, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    T = trk.reshape(-1, 28, 32)
    rng = np.random.default_rng(seed)
    T[:, 4, :] += (2 * np.pi * rng.integers(-2, 3, T[:, 4, :].shape)).astype(np.float32)
    small = rng.random(T[:, 3, :].shape) < 0.3
    T[:, 3, :] = np.where(small, T[:, 3, :] * 1e-3, T[:, 3, :]).astype(np.float32)
    return shape, trk, hits


def test_option_wrap_phi_on_one_layer(orc):
    """One layer with options=1 equals options=0 except phi, which is phi0 - 2 pi n with the n of the formula,
    lands in [-pi, pi] and leaves cos/sin unchanged to FP32 rounding."""
    shape, trk, hits = _options_inputs()
    h1 = np.ascontiguousarray(hits.reshape(-1, 20, 9, 32)[:, :1]).ravel()
    o0 = orc.run_f32(trk, h1, shape.nblk, 1, 32, semantics=1).reshape(-1, 28, 32)
    o1 = orc.run_f32(trk, h1, shape.nblk, 1, 32, semantics=1, options=1).reshape(-1, 28, 32)
    keep = [f for f in range(28) if f != 4]
    assert np.array_equal(o0[:, keep].view(np.uint32), o1[:, keep].view(np.uint32))
    phi0, phi1 = o0[:, 4].astype(np.float64), o1[:, 4].astype(np.float64)
    assert np.all(phi1 >= -np.pi - 1e-6) and np.all(phi1 <= np.pi + 1e-6)
    assert np.any(np.abs(phi0) > np.pi + 0.1)  # the inputs exercise the wrap
    np.testing.assert_allclose(np.cos(phi1), np.cos(phi0), atol=1e-5)
    np.testing.assert_allclose(np.sin(phi1), np.sin(phi0), atol=1e-5)
    n = np.round((phi0 - phi1) / (2 * np.pi))
    np.testing.assert_allclose(phi0 - 2 * np.pi * n, phi1, atol=1e-5)


def test_option_charge_from_ipt(orc):
    """With options=2 a negative updated 1/pt comes out positive with the charge flipped (one layer:
    everything else equal); over 20 chained layers the flipped tracks then propagate with the new charge,
    so they differ from options=0, and unflipped tracks stay identical."""
    shape, trk, hits = _options_inputs()
    h1 = np.ascontiguousarray(hits.reshape(-1, 20, 9, 32)[:, :1]).ravel()
    o0 = orc.run_f32(trk, h1, shape.nblk, 1, 32, semantics=1).reshape(-1, 28, 32)
    o2 = orc.run_f32(trk, h1, shape.nblk, 1, 32, semantics=1, options=2).reshape(-1, 28, 32)
    neg = o0[:, 3] < 0
    assert neg.any() and (~neg).any()
    q0, q2 = o0[:, 27].view(np.int32), o2[:, 27].view(np.int32)
    assert np.array_equal(q2[neg], -q0[neg]) and np.array_equal(q2[~neg], q0[~neg])
    assert np.array_equal(o2[:, 3][neg], -o0[:, 3][neg]) and np.array_equal(o2[:, 3][~neg], o0[:, 3][~neg])
    keep = [f for f in range(27) if f != 3]
    assert np.array_equal(o0[:, keep].view(np.uint32), o2[:, keep].view(np.uint32))
    # reference semantics: every layer restarts from the input charge; the last layer's flip is what comes out
    r2 = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=0, options=2).reshape(-1, 28, 32)
    r0 = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=0).reshape(-1, 28, 32)
    negr = r0[:, 3] < 0
    assert np.array_equal(r2[:, 27].view(np.int32)[negr], -r0[:, 27].view(np.int32)[negr])


def test_option_chi2_per_layer(orc):
    """options=4: chi2 per (layer, track), layer-major.  Pinned three ways: (1) summed over layers in layer
    order (FP32, sequential) it equals the per-track sum bit for bit; (2) in reference semantics layer L
    restarts from the input, so its chi2 equals a one-layer run on hit L alone; (3) with a mask common to
    all tracks, the masked chained run's chi2 at the present layers equals the unmasked run over just those
    layers' hits, and absent layers read 0."""
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(1, 128, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    n, nl = shape.ntrk_total, 20
    for sem in (0, 1):
        _, tot = orc.run_f32(trk, hits, shape.nblk, nl, 32, semantics=sem, chi2=True)
        _, per = orc.run_f32(trk, hits, shape.nblk, nl, 32, semantics=sem, chi2=True, options=4)
        per = per.reshape(nl, n)
        acc = np.zeros(n, np.float32)
        for l in range(nl):
            acc = acc + per[l]
        assert np.array_equal(acc.view(np.uint32), tot.view(np.uint32))
    H = hits.reshape(shape.nblk, nl, 9, 32)
    _, per0 = orc.run_f32(trk, hits, shape.nblk, nl, 32, semantics=0, chi2=True, options=4)
    per0 = per0.reshape(nl, n)
    for L in (0, 7, 19):
        hL = np.ascontiguousarray(H[:, L:L + 1]).ravel()
        _, one = orc.run_f32(trk, hL, shape.nblk, 1, 32, semantics=0, chi2=True)
        assert np.array_equal(one.view(np.uint32), per0[L].view(np.uint32))
    present = [1, 2, 5, 11, 12, 18]
    bits = sum(1 << l for l in present)
    mask = np.full(n, bits, np.uint32)
    _, pm = orc.run_f32(trk, hits, shape.nblk, nl, 32, semantics=1, chi2=True, mask=mask, options=4)
    pm = pm.reshape(nl, n)
    hc = np.ascontiguousarray(H[:, present]).ravel()
    _, pc = orc.run_f32(trk, hc, shape.nblk, len(present), 32, semantics=1, chi2=True, options=4)
    pc = pc.reshape(len(present), n)
    for j, l in enumerate(present):
        assert np.array_equal(pm[l].view(np.uint32), pc[j].view(np.uint32))
    absent = [l for l in range(nl) if l not in present]
    assert np.all(pm[absent] == 0)


@pytest.mark.parametrize("seed", range(4))
def test_transport_preserves_positive_definiteness(orc, seed):
    """J P J^T of a positive-definite P stays positive definite (J is non-singular: unit rows 3, 5 and
    J22 = 1); FP64 lane, random PD covariances around the C0 track, several layers."""
    rng = np.random.default_rng(100 + seed)
    d = json.load(open(os.path.join(GOLDEN, "input_track_c0.json")))
    par = np.array(d["track"]["par"])
    for layer in (0, 9, 19):
        P = _random_pd(rng, 6, 1e-4)
        hit = np.array(d["hits"][21 - layer]["pos"])
        _, c, J = orc.propagate_lane_f64(par, _sym_pack(P), 1, hit, float(orc.kfact_tbb()))
        assert abs(np.linalg.det(J)) > 0
        assert np.linalg.eigvalsh(_sym_full(c)).min() > 0


@pytest.mark.parametrize("seed", range(4))
def test_update_does_not_increase_variances(orc, seed):
    """P' = P - K H P with K = P H^T S^-1 removes a positive semi-definite term: for PD P and PD hit
    covariance no diagonal variance grows, and P' stays PD."""
    rng = np.random.default_rng(200 + seed)
    P = _random_pd(rng, 6, 1e-4)
    Vf = _random_pd(rng, 3, 1e-5)
    V = np.array([Vf[0, 0], Vf[1, 0], Vf[1, 1], Vf[2, 0], Vf[2, 1], Vf[2, 2]])
    par = np.array([-30.0, -20.0, 10.0, 0.3, -2.5, 0.9])
    hit = np.array([-31.0, -19.5, 10.3])
    _, c, chi2 = orc.kalman_lane_f64(par, _sym_pack(P), hit, V)
    Pn = _sym_full(c)
    assert np.all(np.diag(Pn) <= np.diag(P) * (1 + 1e-12))
    assert np.linalg.eigvalsh(Pn).min() > 0 and chi2 >= 0


def test_masked_run_is_dense_run_on_compacted_hits(orc):
    """Chained semantics with a layer mask common to a block equals the unmasked run over just the present
    layers' hits, state and charge bit for bit (absent layers: no propagation, no update)."""
    import p2r_amd
    from p2r_amd import Shape
    shape = Shape(1, 128, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    H = hits.reshape(shape.nblk, 20, 9, 32)
    for present in ([0, 3, 4, 9, 15], [19], list(range(0, 20, 2))):
        mask = np.full(shape.ntrk_total, sum(1 << l for l in present), np.uint32)
        masked = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=1, mask=mask)
        dense = orc.run_f32(trk, np.ascontiguousarray(H[:, present]).ravel(), shape.nblk, len(present), 32,
                            semantics=1)
        assert np.array_equal(_bits(masked), _bits(dense)), present
    none = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=1, mask=np.zeros(shape.ntrk_total, np.uint32))
    assert np.array_equal(_bits(none), _bits(trk))  # no present layer: copied through


def test_reference_driver_harness(orc, ref):
    """oracle/ref_acc_bench.cpp runs the reference's compiled block routines in tbb.cpp's driver loop (the
    CPU baseline).  Built strict, it equals the oracle's reference-semantics run word for word (with the
    acc.v3 kfact), which pins the driver loop; the timed build (-O3 -ffast-math) agrees in charge and
    stays finite."""
    import oracle
    import p2r_amd
    from p2r_amd import Shape
    if not oracle.ref_bench_available():
        pytest.skip("oracle/_ref timing harness not built")
    shape = Shape(2, 640, 20, 32)
    trk, hits = p2r_amd.gen_uniform(shape)
    kf = ref.kfact
    strict = oracle.RefBench("libref_acc_drive.so").run(trk, hits, shape.nblk, 20, 2)
    want = orc.run_f32(trk, hits, shape.nblk, 20, 32, semantics=0, kfact=kf)
    assert np.array_equal(_bits(strict), _bits(want))
    fast = oracle.RefBench().run(trk, hits, shape.nblk, 20, 2)
    f, w = fast.reshape(-1, 28, 32), want.reshape(-1, 28, 32)
    assert np.array_equal(f[:, 27].view(np.int32), w[:, 27].view(np.int32))
    assert np.isfinite(f[:, :27]).all()


def test_errors_vs_fp64_measures_planted_perturbations(orc):
    """The per-quantity error summary (reported by bench.py) on the oracle's own FP64 replay: rounding that
    replay to FP32 gives errors at the FP32 rounding level, and a relative perturbation planted in one
    parameter of every track reads back as that perturbation, in that parameter only."""
    import parity
    import p2r_amd
    from p2r_amd import Shape
    sh = Shape(1, 256, 20, BS)
    trk, hits = p2r_amd.gen_uniform(sh)
    p64, c64, _ = orc.run_f64(trk, hits, sh.nblk, 20, BS, semantics=1)
    buf = np.zeros((sh.nblk, 28, BS), np.float32)
    buf[:, :6, :] = p64.reshape(sh.nblk, BS, 6).transpose(0, 2, 1)
    buf[:, 6:27, :] = c64.reshape(sh.nblk, BS, 21).transpose(0, 2, 1)
    e = parity.errors_vs_fp64(buf.ravel(), BS, p64, c64)
    assert e["tracks"] == 256
    for name in ("x", "y", "z", "ipt", "phi", "theta"):
        assert e[name]["max"] <= 2.0 ** -24 * 1.01, (name, e[name])
    assert e["cov"]["max"] <= 2.0 ** -24 * 1.01
    buf[:, 3, :] *= np.float32(1 + 1e-3)  # ipt
    e2 = parity.errors_vs_fp64(buf.ravel(), BS, p64, c64)
    assert abs(e2["ipt"]["max"] - 1e-3) < 2e-6 and abs(e2["ipt"]["rms"] - 1e-3) < 2e-6
    assert e2["x"] == e["x"] and e2["cov"] == e["cov"]
