"""Benchmark of the p2r propagate-to-radius + Kalman-update path on one MI355X.

Workload (one "step" = one pass of the hot path over one batch): 100 events x
9600 tracks x 20 layers, Matriplex bsize 32, FP32, seeded synthetic inputs
(include/p2r_gen.h), reference layer semantics (every layer restarts from the
input track, as the reference's CPU path does).  Inputs are resident in HBM
before the timed region; each step is one p2r_forward launch.

Prints ONE JSON line.  value = track-layer updates/s (nevts*ntrks*nlayer per
step / time per step).  `roofline` places the kernel against the FP32 vector
peak (the kernel is VALU-bound: 1212 FLOP against 47.2 HBM bytes per update),
`cpu_baseline` times the reference's own CPU kernel (compiled from its sources
into oracle/_ref, OpenMP over blocks) on this host's cores on a bounded sample
of the same workload, or the oracle's restatement where that library is absent.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--semantics reference|chained]
                  [--cpu-seconds S] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "p2r-tests_amd"), os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

# FLOPs per track-layer update of the reference algorithm (propagateToR +
# KalmanUpdate), counted by the oracle's op-counting scalar type: 1275 ops in
# the reference source, of which 63 multiply/add a structural 0 or 1 of the
# Jacobian; the roofline credits the 1212 non-trivial ones.
FLOP_PER_UPDATE = 1212
# FLOPs of the reverse sweep of one track-layer (the adjoint of the update, its chi2 and the propagation),
# counted on the product's own per-lane code by tests/opcount_harness.cpp: 1878 plus 164 for each fixed-point
# iteration whose convergence branch is taken, 2107 on average over the benchmark's tracks.  The reverse's
# recompute of the forward values (checkpointing) is implementation work and not credited.
REVERSE_FLOP_PER_UPDATE = 2107
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector (AMD published)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E (AMD published)
HIT_BYTES = 36             # per track-layer
TRK_BYTES = 112            # per track, read once and written once


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--semantics", default="reference", choices=["reference", "chained"])
    ap.add_argument("--precision", default="exact", choices=["exact", "fast"])
    ap.add_argument("--nevts", type=int, default=100)
    ap.add_argument("--ntrks", type=int, default=9600)
    ap.add_argument("--nlayer", type=int, default=20)
    ap.add_argument("--bsize", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="time only the headline variant")
    ap.add_argument("--no-ragged", action="store_true", help="skip the uneven-batch measurement (500 events, masked layers)")
    ap.add_argument("--ragged-events", type=int, default=500)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the fwd+adjoint measurement (100 x 9600 x 20)")
    ap.add_argument("--no-alignment", action="store_true", help="skip the alignment measurement (1e6 tracks x 40 layers)")
    ap.add_argument("--alignment-tracks", type=int, default=1000000)
    ap.add_argument("--no-parity", action="store_true", help="skip the parity spot check")
    ap.add_argument("--no-autotune", action="store_true", help="time the library's default launch choices")
    ap.add_argument("--fwd-waves", type=int, default=None, choices=[0, 1, 3, 4],
                    help="set the forward kernels' register budget (waves per SIMD) instead of autotuning")
    ap.add_argument("--no-transfers", action="store_true", help="skip the host-to-host (PCIe-inclusive) timing")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (steps times) and exit: the "
                         "reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the p2r path runs on a single MI355X "
              "(events would shard across GPUs with no communication; that path is not launched).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import p2r_amd
    from p2r_amd import Shape

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                          "launches": args.steps}), flush=True)
        return

    shape = Shape(args.nevts, args.ntrks, args.nlayer, args.bsize)
    t0 = time.time()
    trk, hits = p2r_amd.gen_uniform(shape)
    gen_s = time.time() - t0
    dev = torch.device("cuda:0")
    d_trk = torch.from_numpy(trk).to(dev)
    d_hit = torch.from_numpy(hits).to(dev)
    d_out = torch.empty_like(d_trk)
    ctx = p2r_amd.Context(0)
    stream = torch.cuda.current_stream()
    args._shape, args._bufs = shape, (d_trk, d_hit, d_out)

    def step(sem=args.semantics, prec=args.precision):
        ctx.forward(d_trk, d_hit, d_out, shape, sem, stream=stream, precision=prec)

    # launch choices for this shape, measured on these buffers (workgroup size x layers per launch)
    tuning = None
    if args.fwd_waves is not None:
        ctx.set_tuning(fwd_waves=args.fwd_waves)
        tuning = {"chosen": ctx.get_tuning().as_dict(), "candidates": None, "set_by": "--fwd-waves"}
    elif not args.no_autotune:
        rep = ctx.autotune_forward(d_trk, d_hit, d_out, shape, args.semantics, precision=args.precision, reps=5,
                                   stream=stream)
        tuning = {"chosen": ctx.get_tuning().as_dict(), "candidates": rep.as_list()}

    # the GPU's clocks ramp up over the first tens of milliseconds of work: 0.3 s of steps first, then the
    # W warmup steps of the contract
    t_ramp = time.perf_counter()
    while time.perf_counter() - t_ramp < 0.3:
        for _ in range(20):
            step()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_time(e1) / args.steps
    updates = shape.ntrk_total * shape.nlayer
    value = updates / (ms * 1e-3)

    # the other semantics / precision, same inputs (not the headline)
    variants = []
    for sem in (() if args.no_variants else ("reference", "chained")):
        for prec in ("exact", "fast"):
            if (sem, prec) == (args.semantics, args.precision):
                continue
            for _ in range(3):
                step(sem, prec)
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(10):
                step(sem, prec)
            e1.record(stream)
            torch.cuda.synchronize()
            vms = e0.elapsed_time(e1) / 10
            variants.append({"semantics": sem, "precision": prec, "ms_per_step": vms,
                             "value": updates / (vms * 1e-3),
                             "roofline_frac": updates / (vms * 1e-3) * FLOP_PER_UPDATE / 1e12 / PEAK_FP32_TFLOPS})

    # parity spot check of this run's output on the first event against the oracle
    parity_res = None
    try:
        if args.no_parity:
            raise RuntimeError("skipped (--no-parity)")
        import oracle
        import parity
        step()
        torch.cuda.synchronize()
        nb1 = shape.nb
        gpu1 = d_out[: nb1 * 28 * shape.bsize].cpu().numpy()
        orc = oracle.Oracle()
        sem = 0 if args.semantics == "reference" else 1
        t1, h1 = trk[: nb1 * 28 * shape.bsize], hits[: nb1 * shape.nlayer * 9 * shape.bsize]
        if args.precision == "exact":
            cpu1 = orc.run_f32(t1, h1, nb1, shape.nlayer, shape.bsize, semantics=sem, cr_trig=True)
            same = (gpu1.view(np.uint32) == cpu1.view(np.uint32)) | (np.isnan(gpu1) & np.isnan(cpu1))
            parity_res = {"event0_bit_identical_to_oracle": bool(same.all()), "words_differing": int((~same).sum())}
            # error of each quantity against the FP64 replay of the same algorithm, beside that of the
            # restatement built with the reference's CPU flags (-O3 -ffast-math)
            alt1 = oracle.Oracle("liboracle_bench.so").run_f32(t1, h1, nb1, shape.nlayer, shape.bsize, semantics=sem)
            p64, c64, _ = orc.run_f64(t1, h1, nb1, shape.nlayer, shape.bsize, semantics=sem)
            parity_res["event0_rel_err_vs_fp64"] = parity.errors_vs_fp64(gpu1, shape.bsize, p64, c64)
            parity_res["event0_rel_err_vs_fp64_reference_cpu_flags"] = parity.errors_vs_fp64(alt1, shape.bsize, p64,
                                                                                              c64)
            # the 1e-4 bar against the reference's CPU path: the oracle with libm's sinf/cosf is word for word
            # the reference's compiled routines, so this compares this run's GPU output with the reference;
            # beside it the same numbers for the reference's algorithm built with its own flags
            ref1 = orc.run_f32(t1, h1, nb1, shape.nlayer, shape.bsize, semantics=sem)
            v = parity.vs_reference(gpu1, ref1, alt1, shape.bsize)
            for d in (v, v["reference_build_flags"]):
                d.pop("per_field_frac_within_1e-4", None)
            parity_res["event0_vs_reference_cpu"] = v
        else:
            cpu1 = orc.run_f32(t1, h1, nb1, shape.nlayer, shape.bsize, semantics=sem)
            alt1 = oracle.Oracle("liboracle_bench.so").run_f32(t1, h1, nb1, shape.nlayer, shape.bsize, semantics=sem)
            p64, c64, _ = orc.run_f64(t1, h1, nb1, shape.nlayer, shape.bsize, semantics=sem)
            r = parity.compare_statistical(gpu1, cpu1, alt1, shape.bsize, p64, c64)
            parity_res = {"event0_ok": r["ok"], "frac_within_1e-4_of_cpu": r.get("frac_within_rtol_of_cpu"),
                          "frac_within_1e-4_of_cpu_reference_build": r.get("frac_within_rtol_of_cpu_reference_build"),
                          "fields_failing": r["fields_failing"]}
    except Exception as ex:  # the bench line is still printed; parity is covered by the tests
        parity_res = {"error": repr(ex)}

    # the reference's default timing includes the host<->device copies (hip.cu:941-1013): the same batch from
    # pinned host buffers through p2r_forward_streamed; reported beside `value`, never as it
    transfers = None
    if not args.no_transfers:
        ht, hh = d_trk.cpu().pin_memory(), d_hit.cpu().pin_memory()
        ho = torch.empty_like(ht).pin_memory()
        d_in = torch.empty_like(d_trk)
        transfers = {"bytes_per_step": int(2 * ht.numel() * 4 + hh.numel() * 4)}
        for ns in (1, 4):
            ctx.forward_streamed(ht, hh, ho, d_in, d_hit, d_out, shape, args.semantics, precision=args.precision,
                                 nstreams=ns)
            t0 = time.perf_counter()
            for _ in range(3):
                ctx.forward_streamed(ht, hh, ho, d_in, d_hit, d_out, shape, args.semantics,
                                     precision=args.precision, nstreams=ns)
            sec = (time.perf_counter() - t0) / 3
            transfers[f"nstreams_{ns}"] = {"ms": sec * 1e3, "updates_per_s": updates / sec,
                                           "host_device_GBs": transfers["bytes_per_step"] / sec / 1e9}
        del ht, hh, ho, d_in

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_bench(ctx, args,
                                {(v["semantics"], v["precision"]): v["ms_per_step"] for v in variants},
                                (args.semantics, args.precision, ms))
    ragged = None
    if not args.no_ragged:
        del d_trk, d_hit, d_out
        args._bufs = None
        torch.cuda.empty_cache()
        ragged = ragged_bench(ctx, args, {v["precision"]: v for v in variants if v["semantics"] == "chained"},
                              (args.semantics, args.precision, value))

    alignment = None
    if not args.no_alignment:
        torch.cuda.empty_cache()
        alignment = alignment_bench(ctx, args)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(shape, trk, hits, args.semantics, args.cpu_seconds)

    achieved_tflops = value * FLOP_PER_UPDATE / 1e12
    per = ctx.get_tuning().fwd_layers_per_launch
    launches_per_step = 1 if per <= 0 or per >= shape.nlayer else -(-shape.nlayer // per)
    bytes_per_launch = shape.ntrk_total * (2 * TRK_BYTES + HIT_BYTES * shape.nlayer)
    traffic = measured_traffic()
    line = {
        "metric": "track-layer updates/s, fwd and fwd+adjoint, 1 MI355X FP32 vs FP64 oracle",
        "value": value,
        "unit": "track-layer updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded counter-based generator, 100 events x 9600 tracks x 20 layers)",
        "config": {"workload": f"p2r fwd, {args.nevts} events x {args.ntrks} tracks x {args.nlayer} layers, "
                               f"bsize {args.bsize}, {args.semantics} semantics, {args.precision} FP32",
                   "nevts": args.nevts, "ntrks": args.ntrks, "nlayer": args.nlayer, "bsize": args.bsize,
                   "semantics": args.semantics, "precision": args.precision, "pass": "forward"},
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "launches_per_step": launches_per_step,
                     "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "flop_per_update": FLOP_PER_UPDATE,
                     "hbm_algorithmic_GBs": bytes_per_launch / (ms * 1e-3) / 1e9,
                     "hbm_frac": bytes_per_launch / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                     "kernel": "fwd_uniform_kernel"},
        "cpu_baseline": cpu,
        "autotune": tuning,
        "transfer_inclusive": transfers,
        "variants": variants,
        "adjoint": adjoint,
        "ragged": ragged,
        "alignment": alignment,
        "parity": parity_res,
        "wall_s_timed": wall,
        "gen_s": gen_s,
    }
    print(json.dumps(line), flush=True)


def _ramp(fn, seconds=0.3):
    """Calls fn for `seconds` of wall time (the GPU's clocks ramp up over the first tens of ms of work)."""
    import torch
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        fn()
        torch.cuda.synchronize()


def _timed(fn, reps=5, warm=2):
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def adjoint_bench(ctx, args, variant_ms, headline):
    """Fit + adjoint: the headline shape (100 x 9600 x 20), chained fit + reverse-mode gradient of the total chi2 with
    respect to the initial parameters, every hit coordinate and kfact (checkpoints every 4 layers).
    Reported as the time of one fused forward+adjoint launch and its ratio to the plain chained forward."""
    import torch
    shape = args._shape
    d_trk, d_hit, d_out = args._bufs
    n = shape.ntrk_total
    c2 = torch.empty(n, device="cuda")
    gpar = torch.empty(shape.nblk * 6 * shape.bsize, device="cuda")
    ghit = torch.empty(shape.nblk * shape.nlayer * 3 * shape.bsize, device="cuda")
    gk = torch.empty(n, device="cuda")
    res = {"shape": list(shape.tuple()), "checkpoint_stride": 4, "stride_sweep": {},
           "flop_per_update": FLOP_PER_UPDATE + REVERSE_FLOP_PER_UPDATE, "default_grad_precision": "mixed"}
    # the section before this one leaves the GPU idle (host copies): 0.3 s of calls bring its clocks up again
    # before the first timed configuration, as for the headline
    _ramp(lambda: ctx.forward_adjoint(d_trk, d_hit, d_out, shape, grad_precision="f32", checkpoint_stride=4,
                                      chi2=c2, grad_par=gpar, grad_hits=ghit, grad_kfact=gk))
    for prec in ("exact", "fast"):
        fwd_ms = variant_ms.get(("chained", prec))
        if fwd_ms is None and headline[0] == "chained" and headline[1] == prec:
            fwd_ms = headline[2]
        for gp in ("f32", "mixed", "f64"):
            for cs in ((4, 1, 2, 8) if gp != "f64" else (4, 1)):
                ms = _timed(lambda: ctx.forward_adjoint(d_trk, d_hit, d_out, shape, grad_precision=gp, precision=prec,
                                                        checkpoint_stride=cs, chi2=c2, grad_par=gpar, grad_hits=ghit,
                                                        grad_kfact=gk))
                rate = shape.ntrk_total * shape.nlayer / (ms * 1e-3)
                r = {"ms": ms, "updates_per_s": rate,
                     "ratio_to_chained_forward": (ms / fwd_ms) if fwd_ms else None,
                     "roofline_frac": rate * (FLOP_PER_UPDATE + REVERSE_FLOP_PER_UPDATE) / 1e12 / PEAK_FP32_TFLOPS}
                if cs == 4:
                    res[f"{prec}_{gp}"] = r
                else:
                    res["stride_sweep"][f"{prec}_{gp}_cs{cs}"] = r
    return res


def alignment_bench(ctx, args):
    """Alignment workload: 1e6 tracks x 40 layers, hits generated with hidden per-layer rigid offsets
    (dx, dy, dz ~ N(0, 50 um), dphi ~ N(0, 1e-4)) and field scale 1 + N(0, 1e-3); the fit runs at nominal
    geometry and field; fwd+adjoint gives d(sum chi2)/d(160 offsets + field scale), reduced in FP64.
    Reports the time and the relative error of the FP32 / mixed gradients against the FP64 one."""
    import numpy as np
    import torch

    import p2r_amd
    from p2r_amd import Shape
    nl, bs = 40, args.bsize
    ntrks = args.alignment_tracks
    shape = Shape(1, ntrks - ntrks % bs, nl, bs)
    rng = np.random.default_rng(20240905)
    offsets = np.zeros((nl, 4), np.float32)
    offsets[:, :3] = rng.normal(0, 50e-4, (nl, 3))   # 50 um in cm
    offsets[:, 3] = rng.normal(0, 1e-4, nl)
    bscale_truth = float(1 + rng.normal(0, 1e-3))
    t0 = time.time()
    cfg = p2r_amd.gen_config(r0=20.0, dr=2.5, layer_offsets=offsets.ravel(), b_scale_truth=bscale_truth)
    trk, hits = p2r_amd.gen_uniform(shape, cfg)
    gen_s = time.time() - t0
    dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
    del trk, hits
    do = torch.empty_like(dt)
    res = {"tracks": shape.ntrk_total, "nlayer": nl, "gen_s": gen_s, "b_scale_truth": bscale_truth}
    grads = {}
    for gp in ("f64", "mixed", "f32"):
        gg = torch.empty(4 * nl + 1, dtype=torch.float64, device="cuda")
        ms = _timed(lambda: ctx.forward_adjoint(dt, dh, do, shape, grad_precision=gp, precision="exact",
                                                grad_global=gg), reps=3, warm=1)
        grads[gp] = gg.cpu().numpy()
        res[gp] = {"ms": ms, "updates_per_s": shape.ntrk_total * nl / (ms * 1e-3)}
    ref = grads["f64"]
    for gp in ("mixed", "f32"):
        res[gp]["rel_err_vs_f64_norm"] = float(np.linalg.norm(grads[gp] - ref) / np.linalg.norm(ref))
        res[gp]["rel_err_vs_f64_max_component"] = float(np.max(np.abs(grads[gp] - ref) / (np.abs(ref) + 1e-30)))
    g = ref[:4 * nl].reshape(nl, 4)
    # the hits carry the hidden motion, so the correction that undoes it is -offsets; a descent step -grad
    # points along it when grad correlates positively with the offsets
    res["corr_grad_vs_hidden_offsets"] = {k: float(np.corrcoef(g[:, i], offsets[:, i])[0, 1])
                                          for i, k in enumerate(("dx", "dy", "dz", "dphi"))}
    res["d_chi2_d_bscale"] = float(ref[-1])
    return res


def ragged_bench(ctx, args, uniform_chained, headline):
    """Uneven batch: 500 events with lognormal track counts (median 9600, sigma 0.5, clipped to
    [500, 40000]), hit count per track uniform in 5..20 (random subset of the 20 layers), chained
    semantics, one launch replayed from a captured HIP graph.  For each Matriplex width 16/32/64 (inputs
    generated in that width) the autotuner times the schedules (layer order at 64/128/256 lanes per
    workgroup, the LDS pair stream, hit-count-sorted tiles of 1/2/4 rounds); the chosen one is then
    captured and replayed.
    Rate = present-layer updates/s, compared with the uniform batch of the headline shape, chained."""
    import numpy as np
    import torch

    import p2r_amd
    counts = p2r_amd.lognormal_counts(args.ragged_events, median=9600, sigma=0.5, lo=500, hi=40000)
    res = {"events": args.ragged_events, "tracks": int(counts.sum()), "widths": {}}
    best = {}
    stream = torch.cuda.Stream()
    keep = ctx.get_tuning()
    for bs in (16, 32, 64):
        t0 = time.time()
        trk, hits, mask, offs = p2r_amd.gen_ragged(counts, 20, bs, kmin=5, kmax=20)
        gen_s = time.time() - t0
        updates = int(np.unpackbits(mask.view(np.uint8)).sum())
        res["present_updates"] = updates  # the same tracks and layer subsets in every width
        dt, dh = torch.from_numpy(trk).cuda(), torch.from_numpy(hits).cuda()
        dm = torch.from_numpy(mask.view(np.int32)).cuda()
        do = torch.empty_like(dt)
        del trk, hits
        for prec in ("exact", "fast"):
            with torch.cuda.stream(stream):
                rep = ctx.autotune_ragged(dt, dh, do, offs, dm, 20, bs, "chained", precision=prec, reps=3,
                                          stream=stream)
                chosen = ctx.get_tuning()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=stream):
                    ctx.forward_ragged(dt, dh, do, offs, dm, 20, bs, "chained", precision=prec, stream=stream)
                for _ in range(5):  # untimed replays: clocks and caches settle after the autotuner's candidates
                    g.replay()
                stream.synchronize()
                rounds = []
                for _ in range(3):  # median of 3 rounds of 10 replays
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    for _ in range(10):
                        g.replay()
                    e1.record(stream)
                    stream.synchronize()
                    rounds.append(e0.elapsed_time(e1) / 10)
            stream.synchronize()
            ms = sorted(rounds)[1]
            rate = updates / (ms * 1e-3)
            res["widths"][f"{prec}_bs{bs}"] = {
                "graph_ms": ms, "updates_per_s": rate, "slots": int(mask.size), "gen_s": gen_s,
                "schedule": chosen.ragged_schedule, "threads": chosen.ragged_threads,
                "candidates": [{"schedule": e["schedule"], "threads": e["threads"], "ms": e["ms"]}
                               for e in rep.as_list()]}
            if prec not in best or rate > best[prec][0]:
                best[prec] = (rate, ms, bs, chosen.ragged_schedule, chosen.ragged_threads)
            del g
            ctx.set_tuning(keep)
        del dt, dh, dm, do
        torch.cuda.empty_cache()
    for prec, (rate, ms, bs, sched, thr) in best.items():
        uni = uniform_chained.get(prec, {}).get("value")
        if uni is None and headline[0] == "chained" and headline[1] == prec:
            uni = headline[2]
        res[prec] = {"best_bsize": bs, "best_schedule": sched, "best_threads": thr, "ms": ms, "updates_per_s": rate,
                     "uniform_chained_updates_per_s": uni, "fraction_of_uniform": (rate / uni) if uni else None,
                     "roofline_frac": rate * FLOP_PER_UPDATE / 1e12 / PEAK_FP32_TFLOPS}
    return res


def measured_traffic():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", "traffic_fwd_uniform_100x9600x20.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("bytes_per_launch")
    except Exception:
        return None


def host_cpu():
    """What this process may run on: logical CPUs of the host, those in its affinity mask, the thread count
    the environment grants (OMP_NUM_THREADS, the host's share per GPU where one is set), and the CPU model
    as lscpu names it."""
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    granted = int(omp) if omp and omp.isdigit() and int(omp) > 0 else affinity
    model = ""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for l in out.splitlines():
            if l.strip().lower().startswith("model name"):
                model = l.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if not model:
        try:
            for l in open("/proc/cpuinfo"):
                if l.startswith("model name"):
                    model = l.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return {"os_cpu_count": os.cpu_count(), "affinity_cpus": affinity, "omp_num_threads": omp,
            "threads_used": max(1, min(affinity, granted)), "model": model}


def cpu_baseline(shape, trk, hits, semantics, seconds):
    """The reference CPU path on every thread this process is granted, over the whole benchmark batch (all
    events of the same inputs), each leg repeated until `seconds` of wall time.  Kind "reference": the
    reference's own acc.v3 block routines compiled with its CPU flags (-O3 -ffast-math) in tbb.cpp's driver
    loop, OpenMP over Matriplex blocks (oracle/ref_acc_bench.cpp; reference semantics, bsize 32).  Kind
    "port": the oracle's restatement with the same flags, used when that build is absent or for chained
    semantics; its rate is reported beside the reference's either way."""
    import oracle
    from p2r_amd import Shape
    host = host_cpu()
    threads = host["threads_used"]
    nev = shape.nevts
    s = Shape(nev, shape.ntrks, shape.nlayer, shape.bsize)
    t1, h1 = trk[: s.trk_floats], hits[: s.hit_floats]

    def timed(run):
        run()  # warm-up
        reps, t0 = 0, time.perf_counter()
        while True:
            run()
            reps += 1
            el = time.perf_counter() - t0
            if el >= seconds or reps >= 1000:
                return reps, el, reps * s.ntrk_total * s.nlayer / el

    orc = oracle.Oracle("liboracle_bench.so")
    sem = 0 if semantics == "reference" else 1
    preps, pel, prate = timed(lambda: orc.run_f32(t1, h1, s.nblk, s.nlayer, s.bsize, semantics=sem, nthreads=threads))
    kind, reps, el, rate = "port", preps, pel, prate
    how = "the oracle's restatement of tbb.cpp"
    if semantics == "reference" and oracle.ref_bench_available():
        rb = oracle.RefBench()
        if rb.bs == s.bsize:
            reps, el, rate = timed(lambda: rb.run(t1, h1, s.nblk, s.nlayer, threads))
            kind, how = "reference", "the reference's acc.v3.cpp block routines in tbb.cpp's driver loop"
    return {"value": rate, "unit": "track-layer updates/s", "cores": threads, "kind": kind,
            "sample": f"the whole batch, {nev} events x {shape.ntrks} tracks x {shape.nlayer} layers, "
                      f"{reps} repetitions in {el:.1f} s: {how}, OpenMP over Matriplex blocks on {threads} "
                      f"threads, -O3 -ffast-math -march=x86-64-v3",
            "port_value": prate, "port_sample": f"{preps} repetitions in {pel:.1f} s",
            "cpu_model": host["model"], "host": host, "semantics": semantics}


if __name__ == "__main__":
    main()
