"""torch.autograd.Function of the chained p2r fit (thin extension _p2r_torch over include/p2r.h).

    chi2, par_out, cov_out = p2r_amd.autograd.fit(par, cov, q, hit_pos, hit_cov, kfact)

Inputs are torch device tensors in Matriplex order (nblk = number of MPTRK blocks, bs = bsize):

    par      float32 [nblk, 6, bs]            initial x, y, z, 1/pt, phi, theta   (differentiable)
    cov      float32 [nblk, 21, bs]           packed symmetric covariance
    q        int32   [nblk, bs]               charge
    hit_pos  float32 [nblk, nlayer, 3, bs]    hit x, y, z per layer              (differentiable)
    hit_cov  float32 [nblk, nlayer, 6, bs]    hit covariance
    kfact    0-dim tensor                     field constant (tbb.cpp:654 by default; differentiable)

`split_tracks` / `split_hits` cut reference-laid-out MPTRK / MPHIT buffers into these views.
Outputs: chi2 [nblk*bs] (sum over layers of each update's r^T S^-1 r), and the fitted par / cov
(not differentiable: the adjoint kernel differentiates sum chi2 only).

Backward: the forward pass runs p2r_forward_adjoint when a gradient is needed (one fused launch:
fit, checkpoints every `checkpoint_stride` layers, reverse sweep), keeps the per-track gradients
and scales them by the incoming d loss / d chi2 per track, which is exact because each track's
chi2 depends only on that track's inputs.  Alignment parameters and a field scale compose with
plain torch ops outside the function (move hit_pos, divide kfact), and autograd applies the chain
rule.  The reference has no derivatives; this path is new.
"""
import torch

from . import _capi  # noqa: F401  (loads libp2r_amd.so first: the extension links it)

try:
    from . import _p2r_torch as _ext
except ImportError as ex:  # the product path needs the built extension; no fallback
    raise ImportError("p2r_amd._p2r_torch is not built: run `python setup_torch_ext.py build_ext --inplace` in "
                      "p2r-tests_amd/ (after `make`)") from ex

PRECISION = {"exact": 0, "fast": 1}
GRAD_PRECISION = {"f32": 0, "mixed": 1, "f64": 2}


def split_tracks(trk, bsize):
    """MPTRK buffer (float32, nblk*28*bsize words) -> views par [nblk,6,bs], cov [nblk,21,bs], q int32 [nblk,bs]."""
    t = trk.view(-1, 28, bsize)
    return t[:, :6], t[:, 6:27], t[:, 27].view(torch.int32)


def split_hits(hits, nlayer, bsize):
    """MPHIT buffer (float32) -> views hit_pos [nblk,nlayer,3,bs], hit_cov [nblk,nlayer,6,bs]."""
    h = hits.view(-1, nlayer, 9, bsize)
    return h[:, :, :3], h[:, :, 3:]


def default_kfact(device=None):
    return torch.tensor(_ext.default_kfact(), dtype=torch.float32, device=device)


class FitFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, par, hit_pos, kfact, cov, q, hit_cov, precision, grad_precision, checkpoint_stride):
        nblk, _, bs = par.shape
        nlayer = hit_pos.shape[1]
        if cov.shape != (nblk, 21, bs) or q.shape != (nblk, bs) or hit_pos.shape != (nblk, nlayer, 3, bs) or \
                hit_cov.shape != (nblk, nlayer, 6, bs):
            raise ValueError("inconsistent Matriplex shapes: par %s cov %s q %s hit_pos %s hit_cov %s" %
                             (tuple(par.shape), tuple(cov.shape), tuple(q.shape), tuple(hit_pos.shape),
                              tuple(hit_cov.shape)))
        if q.dtype != torch.int32:
            raise TypeError("q must be int32")
        trk = torch.cat([par.float(), cov.float(), q.view(torch.float32).unsqueeze(1)], dim=1).contiguous()
        hits = torch.cat([hit_pos.float(), hit_cov.float()], dim=2).contiguous()
        kf = float(kfact)
        need = any(ctx.needs_input_grad[:3])
        if need:
            out, chi2, gpar, ghit, gk, _ = _ext.forward_adjoint(trk, hits, 1, nblk * bs, nlayer, bs, kf,
                                                                PRECISION[precision], GRAD_PRECISION[grad_precision],
                                                                checkpoint_stride, False)
            ctx.save_for_backward(gpar.view(nblk, 6, bs), ghit.view(nblk, nlayer, 3, bs), gk)
        else:
            out, chi2 = _ext.forward(trk, hits, 1, nblk * bs, nlayer, bs, 1, kf, PRECISION[precision], True)
        o = out.view(nblk, 28, bs)
        par_out, cov_out = o[:, :6], o[:, 6:27]
        ctx.mark_non_differentiable(par_out, cov_out)
        ctx.kfact_device = kfact.device if torch.is_tensor(kfact) else None
        ctx.shape = (nblk, bs)
        return chi2, par_out, cov_out

    @staticmethod
    def backward(ctx, g_chi2, g_par_out, g_cov_out):
        gpar, ghit, gk = ctx.saved_tensors
        nblk, bs = ctx.shape
        w = g_chi2.reshape(nblk, bs).to(gpar.dtype)
        grad_par = gpar * w[:, None, :] if ctx.needs_input_grad[0] else None
        grad_hit = ghit * w[:, None, None, :] if ctx.needs_input_grad[1] else None
        grad_k = None
        if ctx.needs_input_grad[2]:
            grad_k = (gk.double() * g_chi2.double()).sum().float()
            if ctx.kfact_device is not None:
                grad_k = grad_k.to(ctx.kfact_device)
        return grad_par, grad_hit, grad_k, None, None, None, None, None, None


def fit(par, cov, q, hit_pos, hit_cov, kfact=None, precision="exact", grad_precision="mixed", checkpoint_stride=4):
    """Chained fit with a hand-written adjoint (see the module docstring).  Returns (chi2, par_out, cov_out)."""
    if kfact is None:
        kfact = default_kfact()
    if not torch.is_tensor(kfact):
        kfact = torch.tensor(float(kfact), dtype=torch.float32)
    return FitFunction.apply(par, hit_pos, kfact, cov, q, hit_cov, precision, grad_precision, checkpoint_stride)


def rigid_transform(hit_pos, offsets):
    """Per-layer rigid motion of the hits, differentiable in `offsets` [nlayer, 4] = (dx, dy, dz, dphi):
    rotation by dphi about z, then translation -- the convention of grad_global in p2r_forward_adjoint."""
    x, y, z = hit_pos[:, :, 0], hit_pos[:, :, 1], hit_pos[:, :, 2]
    c = torch.cos(offsets[:, 3])[None, :, None]
    s = torch.sin(offsets[:, 3])[None, :, None]
    xr = c * x - s * y + offsets[:, 0][None, :, None]
    yr = s * x + c * y + offsets[:, 1][None, :, None]
    zr = z + offsets[:, 2][None, :, None]
    return torch.stack([xr, yr, zr], dim=2)
