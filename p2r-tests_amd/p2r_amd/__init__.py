"""p2r on MI355X: Python host side of the propagate-to-radius + Kalman-update path.

The compute runs in lib/libp2r_amd.so (hand-written HIP for gfx950) behind the
C-ABI of include/p2r.h; this package binds it with ctypes, passes device
pointers of torch tensors (torch provides device memory and streams) and
mirrors the reference's entry points (cerati/p2r-tests
src/propagate-tor-test_tbb.cpp: propagateToR :656, KalmanUpdate :433, the
driver loop :873-925).
"""
import ctypes as C

import numpy as np

from ._capi import (P2R_CHAINED, P2R_EXACT, P2R_FAST, P2R_GRAD_F32, P2R_GRAD_MIXED, P2R_GRAD_F64, P2R_RAGGED_AUTO, P2R_RAGGED_LAYER_ORDER, P2R_RAGGED_PAIR_STREAM, P2R_OPT_WRAP_PHI, P2R_OPT_CHARGE_FROM_IPT, P2R_OPT_CHI2_PER_LAYER, P2R_ERR_ARG, P2R_ERR_HIP, P2R_ERR_NOMEM, P2R_ERR_SHAPE, P2R_ERR_UNSUPPORTED,
                    P2R_OK, P2R_REFERENCE, P2RError, lib, p2r_gen_config, p2r_params, p2r_shape, p2r_tune_report,
                    p2r_tuning)
from .matriplex import Shape, tracks_to_arrays, hits_block_view  # noqa: F401

SEMANTICS = {"reference": P2R_REFERENCE, "chained": P2R_CHAINED}


def _ptr(t):
    """Device (torch) or host (numpy) buffer -> void*; None -> NULL."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(C.c_void_p)
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


# ---- buffer checks --------------------------------------------------------------------------------------
# The C-ABI takes raw pointers, so a buffer shorter than its shape implies, or a host array where a device
# one is expected, would reach a kernel.  The methods below check what they pass: device buffers must be
# contiguous CUDA tensors on the context's device, of the right dtype, holding at least the elements the
# shape implies.  Sizes are checked only for shapes the library accepts (the same rules as include/p2r.h);
# for the others the library's own status code is reported.
_TORCH_DTYPES = {"f32": "float32", "i32": "int32", "f64": "float64"}


def _check_dev(t, n, kind, device, name):
    if t is None or n <= 0:
        return
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name}: expected a CUDA tensor, got {type(t).__name__}"
                        + ("" if not isinstance(t, torch.Tensor) else f" on {t.device}"))
    if t.device.index != device:
        raise ValueError(f"{name}: tensor on {t.device}, context on cuda:{device}")
    if t.dtype != getattr(torch, _TORCH_DTYPES[kind]):
        raise TypeError(f"{name}: dtype {t.dtype}, expected torch.{_TORCH_DTYPES[kind]}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor is not contiguous")
    if t.numel() < n:
        raise ValueError(f"{name}: {t.numel()} elements, the shape needs {n}")


def _check_host(t, n, name):
    if t is None or n <= 0:
        return
    import torch
    if isinstance(t, torch.Tensor):
        ok = (not t.is_cuda) and t.dtype == torch.float32 and t.is_contiguous()
        size = t.numel()
    elif isinstance(t, np.ndarray):
        ok = t.dtype == np.float32 and t.flags.c_contiguous
        size = t.size
    else:
        raise TypeError(f"{name}: expected a host tensor or array, got {type(t).__name__}")
    if not ok:
        raise TypeError(f"{name}: expected contiguous float32 host memory")
    if size < n:
        raise ValueError(f"{name}: {size} elements, the shape needs {n}")


P2R_MAX_NLAYER = 64  # include/p2r.h
P2R_NOALLOC_SCRATCH_BYTES = 64 << 20  # include/p2r.h: static scratch of p2r_forward_adjoint_noalloc


def _uniform_ok(sh):
    """p2r_forward / p2r_forward_adjoint accept this shape (their host checks)."""
    return (sh.nevts >= 0 and sh.ntrks >= 0 and 1 <= sh.bsize <= 1024 and 1 <= sh.nlayer <= P2R_MAX_NLAYER
            and sh.ntrks % sh.bsize == 0 and sh.nevts * sh.ntrks < (1 << 31))


def _params_ok(p):
    """The entry points accept these parameters (their host checks)."""
    return p.b_scale != 0 and p.precision in (P2R_EXACT, P2R_FAST) and not (p.options & ~7)


def _ragged_ok(offs, nlayer, bsize):
    """p2r_forward_ragged accepts these offsets and sizes (its host checks)."""
    return (len(offs) >= 1 and 1 <= nlayer <= 32 and 1 <= bsize <= 1024 and offs[0] == 0
            and bool(np.all(np.diff(offs) >= 0)) and int(offs[-1]) * bsize < (1 << 31))


PRECISION = {"exact": P2R_EXACT, "fast": P2R_FAST}
GRAD_PRECISION = {"f32": P2R_GRAD_F32, "mixed": P2R_GRAD_MIXED, "f64": P2R_GRAD_F64}


def default_params(precision="exact", b_scale=1.0, options=0):
    """options: bitwise OR of P2R_OPT_WRAP_PHI, P2R_OPT_CHARGE_FROM_IPT, P2R_OPT_CHI2_PER_LAYER (include/p2r.h)."""
    p = lib().p2r_default_params()
    prec = PRECISION[precision] if isinstance(precision, str) else int(precision)
    return p2r_params(p.kfact, float(b_scale), prec, int(options))


class Graph:
    """An instantiated HIP graph of one captured call (p2r_graph_launch / p2r_graph_destroy)."""

    def __init__(self, handle, device):
        self._h = handle
        self.device = device

    def launch(self, stream=None):
        st = lib().p2r_graph_launch(self._h, _stream(stream))
        if st != P2R_OK:
            raise P2RError(st, "p2r_graph_launch")

    def close(self):
        if self._h:
            lib().p2r_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """A device binding (p2r_create / p2r_destroy)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        st = lib().p2r_create(int(device), C.byref(self._h))
        if st != P2R_OK:
            raise P2RError(st, f"p2r_create(device={device})")
        self.device = device

    def close(self):
        if self._h:
            lib().p2r_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != P2R_OK:
            raise P2RError(st, f"{what}: {lib().p2r_last_error(self._h).decode()}")

    def _check_forward(self, trk, hits, out, sh, p, chi2, what):
        if not (_uniform_ok(sh) and _params_ok(p)):
            return
        n = sh.nevts * sh.ntrks
        for t, m, name in ((trk, 28 * n, "trk"), (hits, 9 * n * sh.nlayer, "hits"), (out, 28 * n, "out"),
                           (chi2, n * (sh.nlayer if p.options & P2R_OPT_CHI2_PER_LAYER else 1), "chi2")):
            _check_dev(t, m, "f32", self.device, f"{what}: {name}")

    def _check_ragged(self, trk, hits, out, offs, mask, nlayer, bsize, p, chi2, what):
        if not (_ragged_ok(offs, nlayer, bsize) and _params_ok(p)):
            return
        nslot = int(offs[-1]) * bsize
        for t, m, name in ((trk, 28 * nslot, "trk"), (hits, 9 * nslot * nlayer, "hits"), (out, 28 * nslot, "out"),
                           (chi2, nslot * (nlayer if p.options & P2R_OPT_CHI2_PER_LAYER else 1), "chi2")):
            _check_dev(t, m, "f32", self.device, f"{what}: {name}")
        _check_dev(mask, nslot, "i32", self.device, f"{what}: layer_mask")

    def _check_adjoint(self, trk, hits, out, sh, p, gp, cs, chi2, grad_par, grad_hits, grad_kfact, grad_global,
                       what):
        if not (_uniform_ok(sh) and _params_ok(p) and p.options == 0 and gp in (0, 1, 2)
                and cs in (0, 1, 2, 4, 8)):
            return
        n, L = sh.nevts * sh.ntrks, sh.nlayer
        for t, m, name in ((trk, 28 * n, "trk"), (hits, 9 * n * L, "hits"), (out, 28 * n, "out"), (chi2, n, "chi2"),
                           (grad_par, 6 * n, "grad_par"), (grad_hits, 3 * n * L, "grad_hits"),
                           (grad_kfact, n, "grad_kfact")):
            _check_dev(t, m, "f32", self.device, f"{what}: {name}")
        _check_dev(grad_global, 4 * L + 1, "f64", self.device, f"{what}: grad_global")

    def forward(self, trk, hits, out, shape, semantics="reference", params=None, chi2=None, stream=None,
                precision=None):
        """Whole batch through all layers (reference layout; torch CUDA float32 tensors)."""
        sem = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        self._check_forward(trk, hits, out, sh, p, chi2, "p2r_forward")
        st = lib().p2r_forward(self._h, _ptr(trk), _ptr(hits), _ptr(out), C.byref(sh), sem, C.byref(p), _ptr(chi2),
                               _stream(stream))
        self._check(st, "p2r_forward")

    def forward_streamed(self, host_trk, host_hits, host_out, dev_trk, dev_hits, dev_out, shape,
                         semantics="reference", params=None, precision=None, nstreams=4):
        """Host-to-host forward, chunks over `nstreams` streams (include/p2r.h p2r_forward_streamed); host_*:
        pinned CPU tensors, dev_*: device tensors of the same sizes.  Returns when the output is on the host."""
        sem = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        if _uniform_ok(sh) and 1 <= nstreams <= 16:
            n = sh.nevts * sh.ntrks
            for t, m, name in ((host_trk, 28 * n, "host_trk"), (host_hits, 9 * n * sh.nlayer, "host_hits"),
                               (host_out, 28 * n, "host_out")):
                _check_host(t, m, f"p2r_forward_streamed: {name}")
            self._check_forward(dev_trk, dev_hits, dev_out, sh, p, None, "p2r_forward_streamed")
        st = lib().p2r_forward_streamed(self._h, _ptr(host_trk), _ptr(host_hits), _ptr(host_out), _ptr(dev_trk),
                                        _ptr(dev_hits), _ptr(dev_out), C.byref(sh), sem, C.byref(p), int(nstreams))
        self._check(st, "p2r_forward_streamed")

    def forward_ragged(self, trk, hits, out, blk_offsets, layer_mask, nlayer, bsize, semantics="chained",
                       params=None, chi2=None, stream=None, precision=None, tile_rounds=0):
        """Uneven batch, one persistent launch (include/p2r.h p2r_forward_ragged).  blk_offsets: host int64
        array of nevts+1 block offsets; trk/hits/out/layer_mask/chi2: device tensors."""
        sem = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
        p = params if params is not None else default_params(precision or "exact")
        offs = np.ascontiguousarray(blk_offsets, np.int64)
        self._check_ragged(trk, hits, out, offs, layer_mask, int(nlayer), int(bsize), p, chi2, "p2r_forward_ragged")
        st = lib().p2r_forward_ragged(self._h, _ptr(trk), _ptr(hits), _ptr(out), _ptr(offs), len(offs) - 1,
                                      _ptr(layer_mask), int(nlayer), int(bsize), sem, C.byref(p), _ptr(chi2),
                                      int(tile_rounds), _stream(stream))
        self._check(st, "p2r_forward_ragged")

    def forward_adjoint(self, trk, hits, out, shape, params=None, grad_precision="mixed", checkpoint_stride=0,
                        chi2=None, grad_par=None, grad_hits=None, grad_kfact=None, grad_global=None, stream=None,
                        precision=None):
        """Chained fit plus reverse-mode gradient of the total chi2 (include/p2r.h p2r_forward_adjoint).
        grad_par: MP6F blocks; grad_hits: MP3F per hit block; grad_kfact: per track; grad_global: float64
        [4*nlayer+1] (per-layer rigid offsets dx, dy, dz, dphi, then the field scale)."""
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        gp = GRAD_PRECISION[grad_precision] if isinstance(grad_precision, str) else int(grad_precision)
        self._check_adjoint(trk, hits, out, sh, p, gp, int(checkpoint_stride), chi2, grad_par, grad_hits, grad_kfact,
                            grad_global, "p2r_forward_adjoint")
        st = lib().p2r_forward_adjoint(self._h, _ptr(trk), _ptr(hits), _ptr(out), C.byref(sh), C.byref(p), gp,
                                       int(checkpoint_stride), _ptr(chi2), _ptr(grad_par), _ptr(grad_hits),
                                       _ptr(grad_kfact), _ptr(grad_global), _stream(stream))
        self._check(st, "p2r_forward_adjoint")

    def forward_adjoint_noalloc(self, trk, hits, out, shape, params=None, grad_precision="mixed",
                                checkpoint_stride=0, chi2=None, grad_par=None, grad_hits=None, grad_kfact=None,
                                grad_global=None, stream=None, precision=None):
        """forward_adjoint with its scratch in the library's static device array (include/p2r.h
        p2r_forward_adjoint_noalloc): no allocation; calls on one GPU must not overlap in time."""
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        gp = GRAD_PRECISION[grad_precision] if isinstance(grad_precision, str) else int(grad_precision)
        self._check_adjoint(trk, hits, out, sh, p, gp, int(checkpoint_stride), chi2, grad_par, grad_hits, grad_kfact,
                            grad_global, "p2r_forward_adjoint_noalloc")
        st = lib().p2r_forward_adjoint_noalloc(self._h, _ptr(trk), _ptr(hits), _ptr(out), C.byref(sh), C.byref(p), gp,
                                               int(checkpoint_stride), _ptr(chi2), _ptr(grad_par), _ptr(grad_hits),
                                               _ptr(grad_kfact), _ptr(grad_global), _stream(stream))
        self._check(st, "p2r_forward_adjoint_noalloc")

    # ---- HIP graphs of one call (include/p2r.h p2r_graph_*) ----
    def graph_forward(self, trk, hits, out, shape, semantics="reference", params=None, chi2=None, precision=None):
        """Captures one p2r_forward call (these buffers, this shape, the current tuning) into a Graph."""
        sem = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        self._check_forward(trk, hits, out, sh, p, chi2, "p2r_graph_forward")
        h = C.c_void_p()
        self._check(lib().p2r_graph_forward(self._h, _ptr(trk), _ptr(hits), _ptr(out), C.byref(sh), sem, C.byref(p),
                                            _ptr(chi2), C.byref(h)), "p2r_graph_forward")
        return Graph(h, self.device)

    def graph_forward_ragged(self, trk, hits, out, blk_offsets, layer_mask, nlayer, bsize, semantics="chained",
                             params=None, chi2=None, precision=None, tile_rounds=0):
        sem = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
        p = params if params is not None else default_params(precision or "exact")
        offs = np.ascontiguousarray(blk_offsets, np.int64)
        self._check_ragged(trk, hits, out, offs, layer_mask, int(nlayer), int(bsize), p, chi2,
                           "p2r_graph_forward_ragged")
        h = C.c_void_p()
        self._check(lib().p2r_graph_forward_ragged(self._h, _ptr(trk), _ptr(hits), _ptr(out), _ptr(offs),
                                                   len(offs) - 1, _ptr(layer_mask), int(nlayer), int(bsize), sem,
                                                   C.byref(p), _ptr(chi2), int(tile_rounds), C.byref(h)),
                    "p2r_graph_forward_ragged")
        return Graph(h, self.device)

    def graph_forward_adjoint(self, trk, hits, out, shape, params=None, grad_precision="mixed", checkpoint_stride=0,
                              chi2=None, grad_par=None, grad_hits=None, grad_kfact=None, grad_global=None,
                              precision=None):
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        gp = GRAD_PRECISION[grad_precision] if isinstance(grad_precision, str) else int(grad_precision)
        self._check_adjoint(trk, hits, out, sh, p, gp, int(checkpoint_stride), chi2, grad_par, grad_hits, grad_kfact,
                            grad_global, "p2r_graph_forward_adjoint")
        h = C.c_void_p()
        self._check(lib().p2r_graph_forward_adjoint(self._h, _ptr(trk), _ptr(hits), _ptr(out), C.byref(sh),
                                                    C.byref(p), gp, int(checkpoint_stride), _ptr(chi2),
                                                    _ptr(grad_par), _ptr(grad_hits), _ptr(grad_kfact),
                                                    _ptr(grad_global), C.byref(h)), "p2r_graph_forward_adjoint")
        return Graph(h, self.device)

    # ---- launch choices and autotuners (include/p2r.h p2r_tuning, p2r_autotune_*) ----
    def get_tuning(self):
        t = p2r_tuning()
        self._check(lib().p2r_get_tuning(self._h, C.byref(t)), "p2r_get_tuning")
        return t

    def set_tuning(self, t=None, **kw):
        t = t if t is not None else self.get_tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._check(lib().p2r_set_tuning(self._h, C.byref(t)), "p2r_set_tuning")
        return t

    def autotune_forward(self, trk, hits, out, shape, semantics="reference", params=None, precision=None, reps=5,
                         stream=None):
        """Times fwd_threads x fwd_layers_per_launch on these buffers, keeps the fastest; returns the report."""
        sem = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        self._check_forward(trk, hits, out, sh, p, None, "p2r_autotune_forward")
        r = p2r_tune_report()
        self._check(lib().p2r_autotune_forward(self._h, _ptr(trk), _ptr(hits), _ptr(out), C.byref(sh), sem,
                                               C.byref(p), int(reps), C.byref(r), _stream(stream)),
                    "p2r_autotune_forward")
        return r

    def autotune_ragged(self, trk, hits, out, blk_offsets, layer_mask, nlayer, bsize, semantics="chained",
                        params=None, precision=None, reps=5, stream=None):
        sem = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
        p = params if params is not None else default_params(precision or "exact")
        offs = np.ascontiguousarray(blk_offsets, np.int64)
        self._check_ragged(trk, hits, out, offs, layer_mask, int(nlayer), int(bsize), p, None, "p2r_autotune_ragged")
        r = p2r_tune_report()
        self._check(lib().p2r_autotune_ragged(self._h, _ptr(trk), _ptr(hits), _ptr(out), _ptr(offs), len(offs) - 1,
                                              _ptr(layer_mask), int(nlayer), int(bsize), sem, C.byref(p), int(reps),
                                              C.byref(r), _stream(stream)), "p2r_autotune_ragged")
        return r

    def autotune_adjoint(self, trk, hits, out, shape, params=None, grad_precision="mixed", chi2=None, grad_par=None,
                         grad_hits=None, grad_kfact=None, grad_global=None, precision=None, reps=3, stream=None):
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        gp = GRAD_PRECISION[grad_precision] if isinstance(grad_precision, str) else int(grad_precision)
        self._check_adjoint(trk, hits, out, sh, p, gp, 0, chi2, grad_par, grad_hits, grad_kfact,
                            grad_global, "p2r_autotune_adjoint")
        r = p2r_tune_report()
        self._check(lib().p2r_autotune_adjoint(self._h, _ptr(trk), _ptr(hits), _ptr(out), C.byref(sh), C.byref(p), gp,
                                               _ptr(chi2), _ptr(grad_par), _ptr(grad_hits), _ptr(grad_kfact),
                                               _ptr(grad_global), int(reps), C.byref(r), _stream(stream)),
                    "p2r_autotune_adjoint")
        return r

    def repack(self, src, src_bsize, dst, dst_bsize, nevts, ntrks, nsub, words, stream=None):
        """Matriplex width conversion on the device (MPTRK: nsub=1, words=28; MPHIT: nsub=nlayer, words=9)."""
        total = int(nevts) * int(ntrks) * int(nsub) * int(words)
        if total > 0 and nsub >= 1 and words >= 1:
            _check_dev(src, total, "f32", self.device, "p2r_repack: src")
            _check_dev(dst, total, "f32", self.device, "p2r_repack: dst")
        self._check(lib().p2r_repack(self._h, _ptr(src), int(src_bsize), _ptr(dst), int(dst_bsize), int(nevts),
                                     int(ntrks), int(nsub), int(words), _stream(stream)), "p2r_repack")

    def autotune_width(self, trk, hits, shape, widths=(16, 32, 64), semantics="reference", params=None,
                       precision=None, reps=5, stream=None):
        sem = SEMANTICS[semantics] if isinstance(semantics, str) else int(semantics)
        sh = p2r_shape(*shape.tuple()) if isinstance(shape, Shape) else shape
        p = params if params is not None else default_params(precision or "exact")
        w = np.ascontiguousarray(widths, np.int32)
        self._check_forward(trk, hits, None, sh, p, None, "p2r_autotune_width")
        r = p2r_tune_report()
        self._check(lib().p2r_autotune_width(self._h, _ptr(trk), _ptr(hits), C.byref(sh), sem, C.byref(p), _ptr(w),
                                             len(w), int(reps), C.byref(r), _stream(stream)), "p2r_autotune_width")
        return r

    def propagateToR(self, inErr, inPar, inChg, msP, outErr, outPar, bsize, nblk, trk_stride, hit_stride,
                     kfact=None, stream=None):
        """Mirror of propagateToR (tbb.cpp:656) over nblk blocks; pointers/tensors address field arrays."""
        k = default_params().kfact if kfact is None else kfact
        st = lib().p2r_propagateToR(self._h, _ptr(inErr), _ptr(inPar), _ptr(inChg), _ptr(msP), _ptr(outErr),
                                    _ptr(outPar), bsize, nblk, trk_stride, hit_stride, k, _stream(stream))
        self._check(st, "p2r_propagateToR")

    def KalmanUpdate(self, trkErr, inPar, hitErr, msP, bsize, nblk, trk_stride, hit_stride, stream=None):
        """Mirror of KalmanUpdate (tbb.cpp:433), in place, over nblk blocks."""
        st = lib().p2r_KalmanUpdate(self._h, _ptr(trkErr), _ptr(inPar), _ptr(hitErr), _ptr(msP), bsize, nblk,
                                    trk_stride, hit_stride, _stream(stream))
        self._check(st, "p2r_KalmanUpdate")


# ---- host generators --------------------------------------------------------

def gen_reference(shape):
    """Reference-compatible inputs (tbb.cpp:243-293 + input_track.h) as host float32 arrays."""
    trk = np.zeros(shape.trk_floats, np.float32)
    hits = np.zeros(shape.hit_floats, np.float32)
    st = lib().p2r_gen_reference(_ptr(trk), _ptr(hits), C.byref(p2r_shape(*shape.tuple())))
    if st != P2R_OK:
        raise P2RError(st, "p2r_gen_reference")
    return trk, hits


def reference_hit(layer):
    p = np.zeros(3, np.float32)
    st = lib().p2r_reference_hit(int(layer), _ptr(p))
    if st != P2R_OK:
        raise P2RError(st, "p2r_reference_hit")
    return p


def gen_config(**kw):
    c = lib().p2r_gen_default_config()
    offsets = kw.pop("layer_offsets", None)
    for k, v in kw.items():
        setattr(c, k, v)
    if offsets is not None:
        offsets = np.ascontiguousarray(offsets, np.float32)
        c.layer_offsets = offsets.ctypes.data
        c._keep = offsets  # keep the array alive with the struct
    return c


def gen_uniform(shape, cfg=None):
    """Seeded benchmark inputs (include/p2r_gen.h) as host float32 arrays."""
    cfg = cfg if cfg is not None else gen_config()
    trk = np.zeros(shape.trk_floats, np.float32)
    hits = np.zeros(shape.hit_floats, np.float32)
    st = lib().p2r_gen_uniform(_ptr(trk), _ptr(hits), C.byref(p2r_shape(*shape.tuple())), C.byref(cfg))
    if st != P2R_OK:
        raise P2RError(st, "p2r_gen_uniform")
    return trk, hits


def lognormal_counts(nevts, median=9600, sigma=0.5, lo=500, hi=40000, seed=0x50325221):
    out = np.zeros(nevts, np.int32)
    st = lib().p2r_gen_lognormal_counts(_ptr(out), nevts, median, sigma, lo, hi, seed)
    if st != P2R_OK:
        raise P2RError(st, "p2r_gen_lognormal_counts")
    return out


def ragged_layout(counts, bsize):
    counts = np.ascontiguousarray(counts, np.int32)
    offs = np.zeros(len(counts) + 1, np.int64)
    nblk = C.c_int64()
    st = lib().p2r_gen_ragged_layout(_ptr(counts), len(counts), bsize, _ptr(offs), C.byref(nblk))
    if st != P2R_OK:
        raise P2RError(st, "p2r_gen_ragged_layout")
    return offs, int(nblk.value)


def gen_ragged(counts, nlayer, bsize, kmin=5, kmax=20, cfg=None):
    """Uneven batch: returns (trk, hits, layer_mask, blk_offsets) host arrays."""
    counts = np.ascontiguousarray(counts, np.int32)
    offs, nblk = ragged_layout(counts, bsize)
    cfg = cfg if cfg is not None else gen_config()
    trk = np.zeros(nblk * 28 * bsize, np.float32)
    hits = np.zeros(nblk * nlayer * 9 * bsize, np.float32)
    mask = np.zeros(nblk * bsize, np.uint32)
    st = lib().p2r_gen_ragged(_ptr(trk), _ptr(hits), _ptr(mask), _ptr(counts), len(counts), _ptr(offs), nlayer, bsize,
                              kmin, kmax, C.byref(cfg))
    if st != P2R_OK:
        raise P2RError(st, "p2r_gen_ragged")
    return trk, hits, mask, offs
