"""Autograd surface: ``Fwi.apply(W, vp, vs, rho) -> (taux, tauz)``.

Mirror of PyFWI's ``PyFWI.torchfwi.Fwi`` (src/PyFWI/torchfwi.py:7-91).
- CPU tensors follow the reference exactly: numpy forward modelling, and
  ``W.gradient(adj, parameterization='dv')`` in backward, with gradients
  returned as CPU tensors.
- GPU tensors stay on the device. The plan reads the model and writes the
  records and gradients in place, on torch's current stream.

The backward kernel is selected by ``W.grad_mode``:
  'pyfwi' (default)  the reference's adjoint (drop-in parity)
  'exact'            the exact discrete adjoint.  Without ``W.ckpt_every``, the
                     forward keeps every state when all shots' states fit in
                     80 % of the free device memory (or the propagator keeps
                     every step anyway, chpr = 100) and the adjoint reads them;
                     otherwise the forward states are recomputed from
                     checkpoints: every ``W.ckpt_every`` steps when set, else
                     every ceil(sqrt(nt)) steps with the shots in passes sized
                     to the free device memory (the two-level schedule of
                     pyfwi_gradient).  Both give the same gradient.
"""
import numpy as np
import torch

from . import _native
from .processing import prepare_residual

__all__ = ["Fwi"]


class Fwi(torch.autograd.Function):
    """Forward modelling with an adjoint-state backward.

    Parameters: W (WavePropagator), vp, vs, rho ([nz, nx] tensors).
    Returns the 'taux' and 'tauz' records of ``W.components`` (pressure for
    components=0), shaped by ``inpa['seimogram_shape']``."""

    @staticmethod
    def forward(ctx, W, vp, vs, rho):
        mode = getattr(W, "grad_mode", "pyfwi")
        ctx.W = W
        ctx.mode = mode
        if vp.is_cuda:
            k = getattr(W, "ckpt_every", 0) or _native.CKPT_AUTO
            keep_all = False
            if mode == "exact" and not getattr(W, "ckpt_every", 0):
                free, _ = torch.cuda.mem_get_info(vp.device)
                keep_all = W.keeps_every_step() or \
                    int(W.ns) * int(W.nt) * W.state_bytes() <= 0.8 * free
            if keep_all:
                d = W.forward_modeling_tensor(vp, vs, rho, ckpt_every=0, keep_all=True)
            else:
                d = W.forward_modeling_tensor(vp, vs, rho,
                                              ckpt_every=k if mode == "exact" else 0)
            ctx.ckpt_every = max(k, 0)
            return d["taux"], d["tauz"]
        model = {"vp": vp.detach().cpu().numpy().astype(np.float32),
                 "vs": vs.detach().cpu().numpy().astype(np.float32),
                 "rho": rho.detach().cpu().numpy().astype(np.float32)}
        d = W.forward_modeling(model, show=False)
        return torch.tensor(d["taux"]), torch.tensor(d["tauz"])

    @staticmethod
    def backward(ctx, adj_tx, adj_tz):
        W = ctx.W
        if adj_tx.is_cuda:
            res = {"taux": adj_tx, "tauz": adj_tz}
            g = W.gradient_tensor(res, parameterization="dv", mode=ctx.mode,
                                  ckpt_every=ctx.ckpt_every)
            return None, g["vp"], g["vs"], g["rho"]
        adj_taux = adj_tx.detach().cpu().numpy()
        adj_tauz = adj_tz.detach().cpu().numpy()
        if adj_tauz.ndim == 3:
            nt, nr, ns = adj_taux.shape
            adj_taux = np.reshape(adj_taux, (nt, nr * ns), "F")
            adj_tauz = np.reshape(adj_tauz, (nt, nr * ns), "F")
        adj_src = prepare_residual({"taux": adj_taux, "tauz": adj_tauz}, 1)
        g = W.gradient(adj_src, False, mode=ctx.mode)
        return None, torch.tensor(g["vp"]), torch.tensor(g["vs"]), torch.tensor(g["rho"])
