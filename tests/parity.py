"""Parity criteria between the HIP path and the oracle (shared by the GPU tests and bench).

Three CPU results of the reference algorithm on the same inputs are used:
  c   the oracle's IEEE-strict FP32 restatement (bit-identical to the
      reference's own compiled propagateToR / KalmanUpdate),
  a   the same source built with the reference's own flags (-O3 -ffast-math,
      propagate-tor-test build.py:91) -- another valid FP32 rounding,
  f   the FP64 replay of the same algorithm.
Each output element has a natural scale s: |value| for a parameter,
sqrt(|P_ii P_jj|) for covariance entry (i,j) (floored at 1e-6 of the field's
median scale, for variances that FP32 cancellation drives to 0).

Criterion (written for an FP32 result g of the same inputs):
  1. well-conditioned elements -- those where BOTH CPU builds are within
     RTOL/10 of the FP64 value -- agree with c element by element:
         |g - c| <= RTOL * s                                   (RTOL = 1e-4);
  2. on the remaining (ill-conditioned) elements no element-wise bound is
     meaningful: the Kalman covariance update P - K H P cancels
     catastrophically for precise hits, and two valid FP32 builds of the
     reference (c and a) then disagree by up to ~1e-1 of the scale.  There,
     per field, the distribution of the GPU's error against FP64 must be no
     worse than the CPU path's: RMS and 99th percentile of |g - f|/s at most
     FP32_FACTOR (= 3) times the larger of those of c and a;
  3. the same tracks have a non-finite output as in c (up to tracks where c
     and a already disagree), and the charges are copied unchanged.
"""
import numpy as np

RTOL = 1e-4
FP32_FACTOR = 3.0

_DIAG = [0, 2, 5, 9, 14, 20]
_IJ = [(i, j) for i in range(6) for j in range(i + 1)]
FIELD_NAMES = ["x", "y", "z", "ipt", "phi", "theta"] + [f"c{i}{j}" for i, j in _IJ]


def _arrays(buf, bs):
    t = np.asarray(buf, np.float32).reshape(-1, 28, bs)
    par = t[:, :6, :].transpose(0, 2, 1).reshape(-1, 6).astype(np.float64)
    cov = t[:, 6:27, :].transpose(0, 2, 1).reshape(-1, 21).astype(np.float64)
    q = np.ascontiguousarray(t[:, 27, :]).view(np.int32).reshape(-1)
    return par, cov, q


def scales(par, cov):
    """Per-element scales of parameters (n,6) and covariance (n,21)."""
    sp = np.abs(par)
    d = np.abs(cov[:, _DIAG])
    sc = np.empty_like(cov)
    for e, (i, j) in enumerate(_IJ):
        sc[:, e] = np.sqrt(d[:, i] * d[:, j])
    s = np.hstack([sp, sc])
    with np.errstate(invalid="ignore"):
        med = np.nanmedian(np.where(np.isfinite(s), s, np.nan), axis=0)
    floor = np.where(np.isfinite(med), 1e-6 * med, 0.0)
    s = np.maximum(s, floor)
    return s[:, :6], s[:, 6:]


def compare(gpu_buf, cpu_buf, bs, par64=None, cov64=None, alt_buf=None, valid=None):
    """Error statistics and the verdict of the criterion above (par64/cov64: FP64 replay)."""
    gp, gc, gq = _arrays(gpu_buf, bs)
    cp, cc, cq = _arrays(cpu_buf, bs)
    g, c = np.hstack([gp, gc]), np.hstack([cp, cc])
    a = None
    if alt_buf is not None:
        ap, ac, _ = _arrays(alt_buf, bs)
        a = np.hstack([ap, ac])
    f = np.hstack([par64, cov64]) if par64 is not None else None
    if valid is not None:
        g, c, gq, cq = g[valid], c[valid], gq[valid], cq[valid]
        a = a[valid] if a is not None else None
        f = f[valid] if f is not None else None
    sp, sc = scales(c[:, :6], c[:, 6:])
    s = np.hstack([sp, sc])
    g_fin, c_fin = np.isfinite(g).all(axis=1), np.isfinite(c).all(axis=1)
    chaotic = np.zeros_like(c_fin) if a is None else (np.isfinite(a).all(axis=1) != c_fin)
    res = {"ntracks": int(len(g)), "nonfinite_gpu": int((~g_fin).sum()), "nonfinite_cpu": int((~c_fin).sum()),
           "nan_pattern_equal": bool(np.array_equal(g_fin[~chaotic], c_fin[~chaotic])),
           "q_equal": bool(np.array_equal(gq, cq))}
    both = g_fin & c_fin & np.isfinite(s).all(axis=1)
    if a is not None:
        both &= np.isfinite(a).all(axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.abs(g - c)[both] / s[both]
    ok_direct = rel <= RTOL
    res["elements"] = int(rel.size)
    res["frac_within_rtol_of_cpu"] = float(ok_direct.mean()) if rel.size else 1.0
    res["max_rel_gpu_vs_cpu"] = float(rel.max()) if rel.size else 0.0
    ok = ok_direct
    rms_bad = np.zeros(27, bool)
    if f is not None:
        fb = f[both]
        with np.errstate(invalid="ignore", divide="ignore"):
            e_g = np.abs(g[both] - fb) / s[both]
            e_c = np.abs(c[both] - fb) / s[both]
            e_a = np.abs(a[both] - fb) / s[both] if a is not None else e_c
        well = (e_c <= RTOL / 10) & (e_a <= RTOL / 10)
        ok = ok_direct | ~well
        res["frac_well_conditioned"] = float(well.mean()) if well.size else 1.0
        res["frac_within_rtol_of_cpu_well_conditioned"] = float(ok_direct[well].mean()) if well.any() else 1.0
        if e_g.size:
            ill = ~well
            cnt = ill.sum(axis=0)
            def stat(e, fn):
                return np.array([fn(e[ill[:, k], k]) if cnt[k] else 0.0 for k in range(27)])
            rms = lambda v: float(np.sqrt(np.mean(v ** 2)))
            q99 = lambda v: float(np.quantile(v, 0.99))
            rg, rc, ra = stat(e_g, rms), stat(e_c, rms), stat(e_a, rms)
            qg, qc, qa = stat(e_g, q99), stat(e_c, q99), stat(e_a, q99)
            rms_bad = (rg > FP32_FACTOR * np.maximum(rc, ra) + 1e-7) | (qg > FP32_FACTOR * np.maximum(qc, qa) + 1e-6)
            res["ill_conditioned_fields_worse_than_cpu"] = [FIELD_NAMES[k] for k in np.nonzero(rms_bad)[0]]
            res["per_field_rms_gpu_vs_fp64"] = [float(v) for v in np.sqrt(np.mean(e_g ** 2, axis=0))]
            res["per_field_rms_cpu_vs_fp64"] = [float(v) for v in np.sqrt(np.mean(e_c ** 2, axis=0))]
            res["max_rel_gpu_vs_fp64"] = float(e_g.max())
            res["max_rel_cpu_vs_fp64"] = float(e_c.max())
    res["violations"] = int((~ok).sum())
    res["ok"] = (res["nan_pattern_equal"] and res["q_equal"] and res["violations"] == 0 and not rms_bad.any())
    return res


def errors_vs_fp64(buf, bs, par64, cov64):
    """Per-quantity error of an FP32 result against the FP64 replay, relative to each element's natural
    scale (as above): max, 99th percentile and RMS over the finite tracks for x, y, z, ipt, phi, theta,
    and over all 21 covariance entries together."""
    p, c, _ = _arrays(buf, bs)
    g = np.hstack([p, c])
    f = np.hstack([par64, cov64])
    sp, sc = scales(par64, cov64)
    s = np.hstack([sp, sc])
    ok = np.isfinite(g).all(axis=1) & np.isfinite(f).all(axis=1) & np.isfinite(s).all(axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        e = np.abs(g[ok] - f[ok]) / s[ok]
    out = {"tracks": int(ok.sum())}
    stat = lambda v: {"max": float(v.max()), "p99": float(np.quantile(v, 0.99)), "rms": float(np.sqrt(np.mean(v ** 2)))}
    for k, name in enumerate(FIELD_NAMES[:6]):
        out[name] = stat(e[:, k])
    out["cov"] = stat(e[:, 6:].ravel())
    return out


def compare_statistical(gpu_buf, cpu_buf, alt_buf, bs, par64, cov64):
    """Criterion for an FP32 result computed with a different rounding (fused multiply-adds, hardware
    reciprocal).  The yardstick is the reference's own CPU build a (its build.py flags), which differs from
    the IEEE-strict CPU result c by rounding alone:
      1. per field, the fraction of elements within RTOL of c is at least that of a, minus 2 points;
      2. per field, the 50/99/99.9 % quantiles of |g - c|/s are at most FP32_FACTOR times the larger of
         those of |a - c|/s and |c - f|/s (the spread between valid FP32 roundings, or the CPU path's own
         FP32 error), plus 1e-7;
      3. per field, the 99 and 99.9 % quantiles of |g - f|/s are at most FP32_FACTOR times the larger of
         c's and a's (no less accurate than the reference CPU path against the FP64 replay);
      4. the same tracks are non-finite, except tracks where FP32 arithmetic is near breakdown (an element
         of c differs from the FP64 replay by more than a tenth of its own scale);
      5. charges are unchanged."""
    g = np.hstack(_arrays(gpu_buf, bs)[:2])
    c = np.hstack(_arrays(cpu_buf, bs)[:2])
    a = np.hstack(_arrays(alt_buf, bs)[:2])
    gq, cq = _arrays(gpu_buf, bs)[2], _arrays(cpu_buf, bs)[2]
    f = np.hstack([par64, cov64])
    sp, sc = scales(c[:, :6], c[:, 6:])
    s = np.hstack([sp, sc])
    fin = lambda x: np.isfinite(x).all(axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        err_cf = np.where(np.isfinite(c) & np.isfinite(f), np.abs(c - f) / s, np.inf)
        breakdown = ~fin(c) | (err_cf.max(axis=1) > 0.1)
    res = {"ntracks": int(len(g)), "nonfinite_gpu": int((~fin(g)).sum()), "nonfinite_cpu": int((~fin(c)).sum()),
           "fp32_breakdown_tracks": int(breakdown.sum()),
           "nan_pattern_equal": bool(np.array_equal(fin(g)[~breakdown], fin(c)[~breakdown])),
           "q_equal": bool(np.array_equal(gq, cq))}
    use = fin(g) & fin(c) & fin(a) & ~breakdown
    with np.errstate(invalid="ignore", divide="ignore"):
        d_g = np.abs(g - c)[use] / s[use]
        d_a = np.abs(a - c)[use] / s[use]
        e_g = np.abs(g - f)[use] / s[use]
        e_c = np.abs(c - f)[use] / s[use]
        e_a = np.abs(a - f)[use] / s[use]
    bad = []
    if d_g.size:
        frac_g = (d_g <= RTOL).mean(axis=0)
        frac_a = (d_a <= RTOL).mean(axis=0)
        res["frac_within_rtol_of_cpu"] = float((d_g <= RTOL).mean())
        res["frac_within_rtol_of_cpu_reference_build"] = float((d_a <= RTOL).mean())
        for k in range(27):
            why = []
            if frac_g[k] < frac_a[k] - 0.02:
                why.append("frac")
            for qq in (0.5, 0.99, 0.999):
                yard = max(np.quantile(d_a[:, k], qq), np.quantile(e_c[:, k], qq))
                if np.quantile(d_g[:, k], qq) > FP32_FACTOR * yard + 1e-7:
                    why.append(f"q{qq}")
            for qq in (0.99, 0.999):
                if np.quantile(e_g[:, k], qq) > FP32_FACTOR * max(np.quantile(e_c[:, k], qq),
                                                                  np.quantile(e_a[:, k], qq)) + 1e-7:
                    why.append(f"q{qq}_vs_fp64")
            if why:
                bad.append(f"{FIELD_NAMES[k]}:{'/'.join(why)}")
        res["max_rel_gpu_vs_cpu"] = float(d_g.max())
        res["max_rel_gpu_vs_fp64"] = float(e_g.max())
        res["max_rel_cpu_vs_fp64"] = float(e_c.max())
    res["fields_failing"] = bad
    res["ok"] = res["nan_pattern_equal"] and res["q_equal"] and not bad
    return res


def rel_to(x_buf, ref_buf, bs):
    """|x - ref| / s per element (n, 27): parameters then covariance, s the reference's natural scale
    (scales above), over the tracks finite in both."""
    xp, xc, _ = _arrays(x_buf, bs)
    rp, rc, _ = _arrays(ref_buf, bs)
    sp, sc = scales(rp, rc)
    X, R, S = np.hstack([xp, xc]), np.hstack([rp, rc]), np.hstack([sp, sc])
    ok = np.isfinite(X).all(axis=1) & np.isfinite(R).all(axis=1)
    return np.abs(X[ok] - R[ok]) / S[ok]


def vs_reference(gpu_buf, ref_buf, alt_buf, bs):
    """The element-wise 1e-4 bar against the reference's CPU path: fractions of parameter and covariance
    elements within RTOL of ref_buf, the largest parameter deviation and the worst element; and the same
    numbers for alt_buf (the reference's algorithm as its own build flags compile it), the yardstick of
    what another valid FP32 rounding keeps."""
    def stats(x):
        r = rel_to(x, ref_buf, bs)
        if not r.size:
            return {"tracks": 0}
        i, k = np.unravel_index(int(np.argmax(r)), r.shape)
        return {"tracks": int(len(r)), "params_frac_within_1e-4": float((r[:, :6] <= RTOL).mean()),
                "params_max_rel": float(r[:, :6].max()), "cov_frac_within_1e-4": float((r[:, 6:] <= RTOL).mean()),
                "frac_within_1e-4": float((r <= RTOL).mean()),
                "worst_element": {"field": FIELD_NAMES[k], "rel": float(r[i, k])},
                "per_field_frac_within_1e-4": [float(v) for v in (r <= RTOL).mean(axis=0)]}
    out = stats(gpu_buf)
    if alt_buf is not None:
        out["reference_build_flags"] = stats(alt_buf)
    return out


def summary_fields(res):
    return {k: v for k, v in res.items() if not k.startswith("per_field")}
