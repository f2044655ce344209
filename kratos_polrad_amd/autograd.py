"""torch.autograd bridge: peel-off Stokes images as a differentiable function
of per-cell log-density, albedo and HG asymmetry.

forward : kp_forward (FP64) of the run on kappa_rho = kappa_ref * exp(log_rho)
backward: kp_adjoint (path replay, FP64) with dL/dimage from autograd

The same packet ids are used in both passes (common random numbers), so the
backward pass differentiates exactly the histories the forward pass drew.
Each call keeps its own medium for its backward pass: the engine's medium is
reloaded from it before the replay, so several images in one loss (or other
forward runs between forward and backward) each get the gradient of their
own medium.
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .engine import RunSpec


def polarized_image(engine, source_x, run: RunSpec, kappa_ref, log_rho, albedo, g):
    """Image [n_bands, n_incl, N, N, 3] (I, Q, U) with autograd support.

    kappa_ref: numpy float array (per cell), the extinction at log_rho = 0.
    log_rho, albedo, g: torch float64 tensors (per cell, any device)."""
    import torch

    class _Image(torch.autograd.Function):
        @staticmethod
        def forward(ctx, lr, al, gg):
            kap = (np.asarray(kappa_ref, dtype=np.float64)
                   * np.exp(lr.detach().cpu().numpy())).astype(np.float32)
            alb = al.detach().cpu().numpy().astype(np.float32)
            asym = gg.detach().cpu().numpy().astype(np.float32)
            engine.set_media(kap, alb, asym)
            r = RunSpec(**{**run.__dict__, "precision": _abi.KP_F64})
            out = engine.alloc_outputs(r, J=False, counters=False)
            engine.forward(source_x, r, out)
            ctx.run = r
            ctx.media = (kap, alb, asym)
            ctx.devs = (lr.device, al.device, gg.device)
            return out["image"].to(lr.device)

        @staticmethod
        def backward(ctx, grad_image):
            dev = engine._torch.device("cuda", engine.device)
            w = grad_image.detach().to(device=dev, dtype=torch.float64).contiguous()
            # this call's medium, whatever ran on the engine since its forward
            engine.set_media(*ctx.media)
            gr = engine.adjoint(source_x, ctx.run, w)
            return (gr["d_logrho"].to(ctx.devs[0]), gr["d_albedo"].to(ctx.devs[1]),
                    gr["d_g"].to(ctx.devs[2]))

    return _Image.apply(log_rho, albedo, g)
