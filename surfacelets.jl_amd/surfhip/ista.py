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

Each iteration is one autograd node that saves only its input c_k.  Its
forward and backward both run A through the streaming analysis, one subband
at a time, so besides the saved iterates a step holds one coefficient-sized
buffer and one subband:
  forward:  r = y - A^T c;  z_i = c_i + (A r)_i;  c'_i = soft(z_i, theta_i)
  backward: given g' = dL/dc', recompute z_i as above;
            G_i = (dsoft/dz at z_i)^T g'_i and dL/dtheta_j += <g'_i, dsoft/dtheta>
            over the subbands i of scale j;
            u = A^T G;  dL/dc = G - A u;  dL/dy = u
(z = (I - A A^T) c + A y, so dz/dc = I - A A^T and dz/dy = A).
"""
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


class _ISTAStep(torch.autograd.Function):
    """c' = soft(c + A (y - A^T c), theta[scale(i)]) subband by subband."""

    @staticmethod
    def forward(ctx, c, theta, y, ws, scales):
        ctx.ws, ctx.scales = ws, scales
        ctx.save_for_backward(c, theta, y)
        cn = _coefficients(ws)
        r = y - _ns.synthesize(ws, c)

        def f(i, band):
            z = _band(c, i) + band
            s = scales[i - 1]
            _band(cn, i).copy_(z if s == 0 else soft_threshold(z, theta[s - 1]))

        _ns.analyze(f, ws, r)
        return cn

    @staticmethod
    def backward(ctx, gn):
        c, theta, y = ctx.saved_tensors
        ws, scales = ctx.ws, ctx.scales
        G = _coefficients(ws)
        gtheta = torch.zeros_like(theta)
        r = y - _ns.synthesize(ws, c)

        def f(i, band):
            s = scales[i - 1]
            gi = gn[:, :, :, i - 1]
            if s == 0:
                _band(G, i).copy_(gi)
                return
            with torch.enable_grad():
                z = (_band(c, i) + band).detach().requires_grad_(True)
                th = theta[s - 1].detach().requires_grad_(True)
                gz, gt = torch.autograd.grad(soft_threshold(z, th), (z, th), gi)
            _band(G, i).copy_(gz)
            gtheta[s - 1] += gt

        _ns.analyze(f, ws, r)
        u = _ns.synthesize(ws, G)

        def h(i, band):
            _band(G, i).sub_(band)

        _ns.analyze(h, ws, u)
        return G, gtheta, u, None, None


def ista(ws, y, theta, iters=10):
    """x = A^T c_K after `iters` ISTA steps from c_0 = 0 on the noisy volume y,
    with per-scale thresholds theta (a tensor of length J, differentiable).
    ws: a workspace of a tight NSSurfacelet for y's size and element type."""
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
    c = _coefficients(ws).zero_()
    for _ in range(iters):
        c = _ISTAStep.apply(c, theta, y, ws, scales)
    return _ag.synthesize(ws, c)
