"""Unrolled ISTA denoising with NSSurfacelet, differentiable in the per-scale
thresholds, with recomputation checkpointing and one-subband streaming.

For a tight (Parseval) NSSurfacelet bank A = analyze and A^T = synthesize
(the reference's tight bank: src/transforms/ns_surfacelet.jl:226-238, the
adjoint test at test/items/test_ns_surfacelet.jl:218-237), so the
synthesis-form problem

    min_c  1/2 ||y - A^T c||^2 + sum_j theta_j ||c_j||_1

has the ISTA iteration with unit step

    c_0 = 0,   z_k = c_k + A (y - A^T c_k),   c_{k+1} = soft(z_k, theta),

and the denoised volume is x = A^T c_K.  soft is the soft threshold by
magnitude, theta_j applies to every directional subband of scale j (scales
1..J, coarse to fine, as in subband_layout) and the lowpass is not
thresholded.

The unrolled loop is one autograd node with segment checkpoints: the
forward keeps only the iterates c_s, c_2s, ... (every s = ceil(sqrt(K))
steps, `checkpoint_every`), and the backward walks the segments from the
last to the first, recomputing each iterate c_k from its segment's
checkpoint when step k's backward needs it.  So at most K/s + 1 iterates
(plus one incoming gradient and one work buffer) are resident, against K
when every iterate is saved, for about K s / 2 extra forward steps.  Each
step's forward and backward run A through the streaming analysis, one
subband at a time:
  forward:  r = y - A^T c;  z_i = c_i + (A r)_i;  c'_i = soft(z_i, theta_i)
  backward: given g' = dL/dc', recompute z_i as above;
            G_i = (dsoft/dz at z_i)^T g'_i and dL/dtheta_j += <g'_i, dsoft/dtheta>
            over the subbands i of scale j;
            u = A^T G;  dL/dc = G - A u;  dL/dy = u
(z = (I - A A^T) c + A y, so dz/dc = I - A A^T and dz/dy = A).
"""
import math

import torch

from . import autograd as _ag
from . import ns as _ns
from ._arrays import empty_colmajor
from ._lib import ArgumentError


def soft_threshold(z, theta):
    """sign(z) max(|z| - theta, 0) for real z; z max(1 - theta / |z|, 0) for
    complex z (shrinkage of the magnitude)."""
    if z.is_complex():
        r = torch.abs(z)
        return z * torch.clamp(1 - theta / torch.clamp(r, min=torch.finfo(r.dtype).tiny), min=0)
    return torch.sign(z) * torch.clamp(torch.abs(z) - theta, min=0)


def scale_of_subbands(t, dims):
    """Scale index of each subband in layout order (0 for the lowpass)."""
    return [e.scale for e in _ns.subband_layout(t, dims)]


def _band(c, i):
    return c[:, :, :, i - 1]


def _coefficients(ws):
    return empty_colmajor(tuple(ws.dims) + (ws.nsub,), ws.dtype, ws.device)


def _step(ws, scales, c, theta, y):
    """c' = soft(c + A (y - A^T c), theta[scale(i)]) subband by subband
    (c = None is c_0 = 0)."""
    cn = _coefficients(ws)
    r = y if c is None else y - _ns.synthesize(ws, c)

    def f(i, band):
        z = band if c is None else _band(c, i) + band
        s = scales[i - 1]
        _band(cn, i).copy_(z if s == 0 else soft_threshold(z, theta[s - 1]))

    _ns.analyze(f, ws, r)
    return cn


def _step_backward(ws, scales, c, theta, y, gn, gtheta):
    """(dL/dc, dL/dy) of one step given gn = dL/dc'; adds dL/dtheta into
    gtheta.  c = None is c_0 = 0 (its gradient is not needed)."""
    G = _coefficients(ws)
    r = y if c is None else y - _ns.synthesize(ws, c)

    def f(i, band):
        s = scales[i - 1]
        gi = gn[:, :, :, i - 1]
        if s == 0:
            _band(G, i).copy_(gi)
            return
        with torch.enable_grad():
            z = (band if c is None else _band(c, i) + band).detach().requires_grad_(True)
            th = theta[s - 1].detach().requires_grad_(True)
            gz, gt = torch.autograd.grad(soft_threshold(z, th), (z, th), gi)
        _band(G, i).copy_(gz)
        gtheta[s - 1] += gt

    _ns.analyze(f, ws, r)
    u = _ns.synthesize(ws, G)
    if c is None:
        return None, u

    def h(i, band):
        _band(G, i).sub_(band)

    _ns.analyze(h, ws, u)
    return G, u


class _ISTA(torch.autograd.Function):
    """c_K of `iters` ISTA steps from c_0 = 0, differentiable in theta and y,
    with segment checkpoints every `every` steps."""

    @staticmethod
    def forward(ctx, theta, y, ws, scales, iters, every):
        ctx.ws, ctx.scales, ctx.iters, ctx.every = ws, scales, iters, every
        ctx.save_for_backward(theta, y)
        ckpt = {0: None}
        c = None
        for k in range(iters):
            if k and k % every == 0:
                ckpt[k] = c
            c = _step(ws, scales, c, theta, y)
        ctx.ckpt = ckpt
        return c

    @staticmethod
    def backward(ctx, gc):
        theta, y = ctx.saved_tensors
        ws, scales, iters, every = ctx.ws, ctx.scales, ctx.iters, ctx.every
        gtheta = torch.zeros_like(theta)
        gy = torch.zeros_like(y)
        for b in reversed(range(0, iters, every)):
            e = min(b + every, iters)
            for k in reversed(range(b, e)):
                # c_k recomputed from the segment's checkpoint c_b.
                c = ctx.ckpt[b]
                for _ in range(b, k):
                    c = _step(ws, scales, c, theta, y)
                gc, u = _step_backward(ws, scales, c, theta, y, gc, gtheta)
                gy += u
                del c
        ctx.ckpt = None
        return gtheta, gy, None, None, None, None


def ista(ws, y, theta, iters=10, checkpoint_every=None):
    """x = A^T c_K after `iters` ISTA steps from c_0 = 0 on the noisy volume y,
    with per-scale thresholds theta (a tensor of length J, differentiable).
    ws: a workspace of a tight NSSurfacelet for y's size and element type.
    checkpoint_every: steps per checkpoint segment (default ceil(sqrt(iters));
    1 keeps every iterate)."""
    if not isinstance(ws, _ns.NSSurfaceletWorkspace):
        raise ArgumentError("ista needs an NSSurfacelet workspace")
    t = ws.transform
    if not t.tight:
        raise ArgumentError("ista uses A^T = synthesize, which needs a tight NSSurfacelet bank")
    if iters < 1:
        raise ArgumentError(f"iters must be ≥ 1, got {iters}")
    if theta.dim() != 1 or theta.numel() != t.scales:
        raise ArgumentError(f"theta needs one threshold per scale ({t.scales}), got shape {tuple(theta.shape)}")
    if tuple(y.shape) != tuple(ws.dims):
        raise ArgumentError(f"volume of size {tuple(y.shape)}, the workspace is for {tuple(ws.dims)}")
    scales = scale_of_subbands(t, ws.dims)
    y = y.to(ws.dtype)
    theta = theta.to(device=ws.device, dtype=torch.float32)
    every = checkpoint_every if checkpoint_every is not None else max(1, math.ceil(math.sqrt(iters)))
    if every < 1:
        raise ArgumentError(f"checkpoint_every must be ≥ 1, got {every}")
    c = _ISTA.apply(theta, y, ws, scales, iters, every)
    return _ag.synthesize(ws, c)
