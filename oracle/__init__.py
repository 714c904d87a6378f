"""ctypes access to the p2r oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (p2r_oracle.hpp / p2r_oracle.cpp) is a CPU restatement of the
reference algorithm (cerati/p2r-tests, src/propagate-tor-test_tbb.cpp) used to
check the HIP product path.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module.

`ref()` loads the reference's own code compiled from its sources by
oracle/Makefile into oracle/_ref/ (absent when no reference checkout was
available at build time).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def build():
    """Compile liboracle*.so and, when the reference checkout exists, _ref/."""
    subprocess.run(["make", "-C", HERE, "-j4", "all"], check=True, stdout=subprocess.DEVNULL)


class Oracle:
    def __init__(self, name="liboracle.so"):
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run `make -C oracle`)")
        L = self.lib = C.CDLL(path)
        vp, i, l, f, d = C.c_void_p, C.c_int, C.c_long, C.c_float, C.c_double
        L.orc_propagateToR_block.argtypes = [i, f, vp, vp, vp, vp, vp, vp]
        L.orc_KalmanUpdate_block.argtypes = [i, vp, vp, vp, vp]
        L.orc_propagate_lane_f64.argtypes = [vp, vp, i, vp, d, vp, vp, vp]
        L.orc_kalman_lane_f64.argtypes = [vp, vp, vp, vp, vp]
        L.orc_run_f32.argtypes = [vp, vp, vp, l, i, i, i, f, i, vp, vp, i]
        L.orc_run_f32.restype = i
        L.orc_run_f64.argtypes = [vp, vp, vp, vp, vp, l, i, i, i, d, i, vp, i]
        L.orc_run_f64.restype = i
        L.orc_kfact_tbb.restype = f
        L.orc_gen_c0.argtypes = [vp, vp, i, i, i, i]
        L.orc_gen_c0.restype = i
        L.orc_c0_hit.argtypes = [i, vp]
        L.orc_rand_prefix.argtypes = [C.c_uint, i, vp]
        L.orc_summary.argtypes = [vp, i, i, i, f, f, f, vp, vp]
        L.orc_opcount.argtypes = [vp]
        L.orc_set_trig_mode.argtypes = [i]
        L.orc_chi2_f64.argtypes = [vp, vp, i, vp, vp, i, d]
        L.orc_chi2_f64.restype = d
        L.orc_chi2_grad_dual.argtypes = [vp, vp, i, vp, i, d, vp, vp]

    # -- block-level mirrors of the reference entry points --
    def propagate_block(self, bs, kfact, inErr, inPar, inChg, msP):
        outErr = np.zeros(21 * bs, np.float32)
        outPar = np.zeros(6 * bs, np.float32)
        self.lib.orc_propagateToR_block(bs, kfact, _ptr(inErr), _ptr(inPar), _ptr(inChg), _ptr(msP),
                                        _ptr(outErr), _ptr(outPar))
        return outErr, outPar

    def kalman_block(self, bs, trkErr, inPar, hitErr, msP):
        trkErr = trkErr.copy()
        inPar = inPar.copy()
        self.lib.orc_KalmanUpdate_block(bs, _ptr(trkErr), _ptr(inPar), _ptr(hitErr), _ptr(msP))
        return trkErr, inPar

    def propagate_lane_f64(self, par, cov, q, hit, kfact):
        po, co, J = np.zeros(6), np.zeros(21), np.zeros(36)
        self.lib.orc_propagate_lane_f64(_ptr(np.ascontiguousarray(par, np.float64)),
                                        _ptr(np.ascontiguousarray(cov, np.float64)), int(q),
                                        _ptr(np.ascontiguousarray(hit, np.float64)), float(kfact),
                                        _ptr(po), _ptr(co), _ptr(J))
        return po, co, J.reshape(6, 6)

    def kalman_lane_f64(self, par, cov, hitpos, hitcov):
        par = np.array(par, np.float64)
        cov = np.array(cov, np.float64)
        c2 = np.zeros(1)
        self.lib.orc_kalman_lane_f64(_ptr(par), _ptr(cov), _ptr(np.ascontiguousarray(hitpos, np.float64)),
                                     _ptr(np.ascontiguousarray(hitcov, np.float64)), _ptr(c2))
        return par, cov, float(c2[0])

    # -- whole batches in the reference Matriplex layout --
    def run_f32(self, trk, hits, nblk, nlayer, bs, semantics=0, kfact=None, nthreads=0, chi2=False, mask=None,
                cr_trig=False, options=0):
        """cr_trig: sin/cos correctly rounded (FP64 rounded to float) instead of libm's sinf/cosf.
        options: 1 = wrap phi after each update, 2 = charge from the sign of 1/pt, 4 = chi2 per layer
        ([nlayer][nblk*bs], 0 for absent layers) instead of the per-track sum (see run_block)."""
        kfact = self.kfact_tbb() if kfact is None else kfact
        self.lib.orc_set_trig_mode(1 if cr_trig else 0)
        out = np.zeros_like(trk)
        c2 = np.zeros(nblk * bs * (nlayer if options & 4 else 1), np.float32) if chi2 else None
        m = None if mask is None else np.ascontiguousarray(mask, np.uint32)
        rc = self.lib.orc_run_f32(_ptr(trk), _ptr(hits), _ptr(out), nblk, nlayer, bs, semantics, kfact, nthreads,
                                  _ptr(c2), _ptr(m), int(options))
        self.lib.orc_set_trig_mode(0)
        assert rc == 0
        return (out, c2) if chi2 else out

    def run_f64(self, trk, hits, nblk, nlayer, bs, semantics=0, kfact=None, nthreads=0, mask=None, options=0):
        kfact = float(self.kfact_tbb()) if kfact is None else kfact
        n = nblk * bs
        par, cov, c2 = np.zeros((n, 6)), np.zeros((n, 21)), np.zeros(n)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint32)
        rc = self.lib.orc_run_f64(_ptr(trk), _ptr(hits), _ptr(par), _ptr(cov), _ptr(c2), nblk, nlayer, bs,
                                  semantics, kfact, nthreads, _ptr(m), int(options))
        assert rc == 0
        return par, cov, c2

    # -- gradients of the chained fit's total chi2 (one track) --
    def chi2_f64(self, par, cov, q, hitpos, hits_aos, nlayer, kfact=None):
        kfact = float(self.kfact_tbb()) if kfact is None else kfact
        return self.lib.orc_chi2_f64(_ptr(np.ascontiguousarray(par, np.float64)),
                                     _ptr(np.ascontiguousarray(cov, np.float32)), int(q),
                                     _ptr(np.ascontiguousarray(hitpos, np.float64)),
                                     _ptr(np.ascontiguousarray(hits_aos, np.float32)), nlayer, kfact)

    def chi2_grad_dual(self, par, cov, q, hits_aos, nlayer, kfact=None):
        """FP64 forward-mode gradient: [d/dpar (6), d/dhit (3 per layer), d/dkfact], and the chi2."""
        kfact = float(self.kfact_tbb()) if kfact is None else kfact
        g = np.zeros(6 + 3 * nlayer + 1)
        c2 = np.zeros(1)
        self.lib.orc_chi2_grad_dual(_ptr(np.ascontiguousarray(par, np.float32)),
                                    _ptr(np.ascontiguousarray(cov, np.float32)), int(q),
                                    _ptr(np.ascontiguousarray(hits_aos, np.float32)), nlayer, kfact, _ptr(g), _ptr(c2))
        return g, float(c2[0])

    def kfact_tbb(self):
        return self.lib.orc_kfact_tbb()

    def gen_c0(self, nevts, ntrks, nlayer, bs):
        nb = ntrks // bs
        trk = np.zeros(nevts * nb * 28 * bs, np.float32)
        hits = np.zeros(nlayer * nevts * nb * 9 * bs, np.float32)
        assert self.lib.orc_gen_c0(_ptr(trk), _ptr(hits), nevts, ntrks, nlayer, bs) == 0
        return trk, hits

    def c0_hit(self, layer):
        p = np.zeros(3, np.float32)
        self.lib.orc_c0_hit(layer, _ptr(p))
        return p

    def rand_prefix(self, seed, n):
        out = np.zeros(n, np.int32)
        self.lib.orc_rand_prefix(seed, n, _ptr(out))
        return out

    def summary(self, out, nevts, ntrks, bs, hit):
        stats = np.zeros(22)
        counts = np.zeros(2, np.int32)
        self.lib.orc_summary(_ptr(out), nevts, ntrks, bs, float(hit[0]), float(hit[1]), float(hit[2]),
                             _ptr(stats), _ptr(counts))
        return stats, counts

    def opcount(self):
        out = np.zeros(7, np.int64)
        self.lib.orc_opcount(_ptr(out))
        keys = ["add", "mul", "div", "sqrt", "trig", "trivial", "chi2"]
        return dict(zip(keys, (int(v) for v in out)))


SUMMARY_KEYS = ["x_avg", "x_std_over_avg", "y_avg", "y_std_over_avg", "z_avg", "z_std_over_avg", "r_avg",
                "r_std_over_avgz", "dx_avg", "dx_std", "dy_avg", "dy_std", "dz_avg", "dz_std", "dr_avg", "dr_std",
                "pt_avg", "phi_avg", "theta_avg", "pt_rms", "phi_rms", "theta_rms"]


def summary_lines(stats, counts):
    """The reference's printed summary (tbb.cpp:1036-1048), same formats."""
    s = stats
    return [
        "track x avg=%f std/avg=%f" % (s[0], s[1]),
        "track y avg=%f std/avg=%f" % (s[2], s[3]),
        "track z avg=%f std/avg=%f" % (s[4], s[5]),
        "track r avg=%f std/avg=%f" % (s[6], s[7]),
        "track dx/x avg=%f std=%f" % (s[8], s[9]),
        "track dy/y avg=%f std=%f" % (s[10], s[11]),
        "track dz/z avg=%f std=%f" % (s[12], s[13]),
        "track dr/r avg=%f std=%f" % (s[14], s[15]),
        "track pt avg=%f" % s[16],
        "track phi avg=%f" % s[17],
        "track theta avg=%f" % s[18],
        "number of tracks with nans=%d" % counts[0],
        "number of tracks failed=%d" % counts[1],
    ]


class Ref:
    """The reference's own compiled routines (oracle/_ref/)."""

    def __init__(self):
        d = os.path.join(HERE, "_ref")
        self.acc = C.CDLL(os.path.join(d, "libref_acc.so"))
        self.pstl = C.CDLL(os.path.join(d, "libref_pstl.so"))
        self.pstl_c0 = os.path.join(d, "pstl_c0")
        vp = C.c_void_p
        self.acc.ref_acc_kfact.restype = C.c_float
        self.acc.ref_acc_propagateToR.argtypes = [vp] * 6
        self.acc.ref_acc_KalmanUpdate.argtypes = [vp] * 4
        self.acc.ref_acc_MultHelixProp.argtypes = [vp] * 3
        self.acc.ref_acc_MultHelixPropTransp.argtypes = [vp] * 3
        self.pstl.ref_pstl_generate.argtypes = [vp, vp]
        self.bs = self.acc.ref_acc_bsize()
        self.kfact = self.acc.ref_acc_kfact()

    def propagate_block(self, inErr, inPar, inChg, msP):
        bs = self.bs
        outErr = np.zeros(21 * bs, np.float32)
        outPar = np.zeros(6 * bs, np.float32)
        self.acc.ref_acc_propagateToR(_ptr(inErr), _ptr(inPar), _ptr(inChg), _ptr(msP), _ptr(outErr), _ptr(outPar))
        return outErr, outPar

    def kalman_block(self, trkErr, inPar, hitErr, msP):
        trkErr = trkErr.copy()
        inPar = inPar.copy()
        self.acc.ref_acc_KalmanUpdate(_ptr(trkErr), _ptr(inPar), _ptr(hitErr), _ptr(msP))
        return trkErr, inPar

    def pstl_generate(self):
        L = self.pstl
        bs, ntrks, nevts, nlayer = L.ref_pstl_bsize(), L.ref_pstl_ntrks(), L.ref_pstl_nevts(), L.ref_pstl_nlayer()
        nb = ntrks // bs
        trk = np.zeros(nevts * nb * 28 * bs, np.float32)
        hits = np.zeros(nlayer * nevts * nb * 9 * bs, np.float32)
        L.ref_pstl_generate(_ptr(trk), _ptr(hits))
        return (nevts, ntrks, nlayer, bs), trk, hits

    def pstl_c0_summary(self):
        """Run the reference program (10 evts x 9600 trks x 20 layers) and return its summary lines."""
        out = subprocess.run([self.pstl_c0], check=True, capture_output=True, text=True).stdout
        return [l.strip() for l in out.splitlines()
                if (l.startswith("track ") and not l.startswith("track in ")) or l.startswith("number of")]


def ref_available():
    return all(os.path.exists(os.path.join(HERE, "_ref", n)) for n in ("libref_acc.so", "libref_pstl.so", "pstl_c0"))


class RefBench:
    """The reference's acc.v3 block routines built with its CPU flags and driven like tbb.cpp's loop
    (oracle/ref_acc_bench.cpp, output _ref/libref_acc_bench.so): the CPU baseline of kind "reference"."""

    NAME = "libref_acc_bench.so"

    def __init__(self, name=None):
        self.lib = C.CDLL(os.path.join(HERE, "_ref", name or self.NAME))
        self.lib.ref_acc_bench_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_int, C.c_int]
        self.bs = self.lib.ref_acc_bench_bsize()

    def run(self, trk, hits, nblk, nlayer, nthreads=0):
        out = np.zeros_like(trk)
        assert self.lib.ref_acc_bench_run(_ptr(trk), _ptr(hits), _ptr(out), nblk, nlayer, nthreads) == 0
        return out


def ref_bench_available():
    return os.path.exists(os.path.join(HERE, "_ref", RefBench.NAME))
